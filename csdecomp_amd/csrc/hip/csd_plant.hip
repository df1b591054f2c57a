// csd_plant.hip -- compiles a csd_plant_desc (the content of the reference's
// MinimalPlant, plant/minimal_plant.h:16-39) into the device image the
// kernels read, and implements the library-level C-ABI entry points.
//
// Compilation folds every fixed joint into the rigid body of its nearest
// non-fixed ancestor (docs/mainpage.dox:71-73 lists this as a TODO of the
// reference), so FK does one 3x4 compose per moving joint. Geometry attached
// to world-fixed links is transformed to the world frame once, in double.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "csd_internal.h"

namespace csd {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? CSD_ERR_OUT_OF_MEMORY : CSD_ERR_HIP;
}

namespace {

// 3x4 row-major affine transform in double
struct Aff {
  double m[12];
  static Aff identity() {
    Aff a;
    memset(a.m, 0, sizeof(a.m));
    a.m[0] = a.m[5] = a.m[10] = 1.0;
    return a;
  }
  static Aff from_colmajor4(const float *X) {
    Aff a;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) a.m[r * 4 + c] = X[c * 4 + r];
    return a;
  }
  Aff operator*(const Aff &o) const {
    Aff c;
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 4; ++k) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += m[r * 4 + j] * o.m[j * 4 + k];
        if (k == 3) s += m[r * 4 + 3];
        c.m[r * 4 + k] = s;
      }
    }
    return c;
  }
  void to_float(float *out) const {
    for (int k = 0; k < 12; ++k) out[k] = (float)m[k];
  }
};

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

// rank used to put pairs in canonical (sphere, box, capsule) order
int type_rank(int t) {
  switch (t) {
    case CSD_SPHERE: return 0;
    case CSD_BOX: return 1;
    case CSD_CAPSULE: return 2;
    default: return 3;
  }
}

// signed distance in double between two world-posed geometries; used for
// pairs whose relative pose does not depend on q (both world-fixed, or both
// on one rigid body). Capsule cases are bounded conservatively.
double static_pair_sd(int ta, const Aff &A, const double *da, int tb,
                      const Aff &B, const double *db) {
  auto pos = [](const Aff &X, double *p) {
    p[0] = X.m[3];
    p[1] = X.m[7];
    p[2] = X.m[11];
  };
  double pa[3], pb[3];
  pos(A, pa);
  pos(B, pb);
  if (ta == CSD_SPHERE && tb == CSD_SPHERE) {
    double d = sqrt((pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                    (pa[2] - pb[2]) * (pa[2] - pb[2]));
    return d - da[0] - db[0];
  }
  if ((ta == CSD_SPHERE && tb == CSD_BOX) || (ta == CSD_BOX && tb == CSD_SPHERE)) {
    const Aff &X = ta == CSD_BOX ? A : B;
    const double *h = ta == CSD_BOX ? da : db;
    const double *c = ta == CSD_BOX ? pb : pa;
    double r = ta == CSD_BOX ? db[0] : da[0];
    double d[3] = {c[0] - X.m[3], c[1] - X.m[7], c[2] - X.m[11]}, l[3];
    for (int i = 0; i < 3; ++i)
      l[i] = X.m[0 * 4 + i] * d[0] + X.m[1 * 4 + i] * d[1] + X.m[2 * 4 + i] * d[2];
    double o2 = 0, in = -1e300;
    bool out = false;
    for (int i = 0; i < 3; ++i) {
      double e = fabs(l[i]) - h[i];
      if (e > 0) {
        o2 += e * e;
        out = true;
      }
      in = std::max(in, e);
    }
    return (out ? sqrt(o2) : in) - r;
  }
  if (ta == CSD_BOX && tb == CSD_BOX) {
    double ax[15][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        ax[i][k] = A.m[k * 4 + i];
        ax[3 + i][k] = B.m[k * 4 + i];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double *u = ax[i], *v = ax[3 + j];
        double *c = ax[6 + 3 * i + j];
        c[0] = u[1] * v[2] - u[2] * v[1];
        c[1] = u[2] * v[0] - u[0] * v[2];
        c[2] = u[0] * v[1] - u[1] * v[0];
      }
    double best = -1e300;
    for (int i = 0; i < 15; ++i) {
      double n = sqrt(ax[i][0] * ax[i][0] + ax[i][1] * ax[i][1] + ax[i][2] * ax[i][2]);
      if (n < 1e-9) continue;
      double u[3] = {ax[i][0] / n, ax[i][1] / n, ax[i][2] / n};
      double lhs = fabs((pa[0] - pb[0]) * u[0] + (pa[1] - pb[1]) * u[1] + (pa[2] - pb[2]) * u[2]);
      double rhs = 0;
      for (int k = 0; k < 3; ++k) {
        rhs += da[k] * fabs(ax[k][0] * u[0] + ax[k][1] * u[1] + ax[k][2] * u[2]);
        rhs += db[k] * fabs(ax[3 + k][0] * u[0] + ax[3 + k][1] * u[1] + ax[3 + k][2] * u[2]);
      }
      best = std::max(best, lhs - rhs);
    }
    return best;
  }
  // capsule involved: bound with the enclosing spheres of the two geometries
  auto bound = [](int t, const double *d) {
    if (t == CSD_SPHERE) return d[0];
    if (t == CSD_CAPSULE) return d[0] + d[1];
    return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  };
  double dist = sqrt((pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                     (pa[2] - pb[2]) * (pa[2] - pb[2]));
  double s = dist - bound(ta, da) - bound(tb, db);
  return s > 0 ? s : -1.0;  // treat as colliding unless provably apart
}

struct Builder {
  std::vector<char> bytes;
  template <class T>
  size_t add(const T *data, size_t count) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + sizeof(T) * std::max<size_t>(count, 1), 0);
    if (count) memcpy(bytes.data() + off, data, sizeof(T) * count);
    return off;
  }
};

}  // namespace
}  // namespace csd

using namespace csd;

extern "C" {

const char *csd_status_string(int status) {
  switch (status) {
    case CSD_OK: return "ok";
    case CSD_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CSD_ERR_CAPACITY: return "capacity exceeded";
    case CSD_ERR_HIP: return "HIP runtime error";
    case CSD_ERR_NO_DEVICE: return "no gfx950 device";
    case CSD_ERR_UNSUPPORTED: return "unsupported";
    case CSD_ERR_OUT_OF_MEMORY: return "out of device memory";
    default: return "unknown status";
  }
}

const char *csd_last_error(void) { return g_last_error.c_str(); }

int csd_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(CSD_ERR_INVALID_ARGUMENT, "null pointer");
  *ptr = nullptr;
  if (bytes == 0) return CSD_OK;
  CSD_HIP(hipMalloc(ptr, bytes));
  return CSD_OK;
}
int csd_free(void *ptr) {
  if (ptr) CSD_HIP(hipFree(ptr));
  return CSD_OK;
}
int csd_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return CSD_OK;
  CSD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return CSD_OK;
}
int csd_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return CSD_OK;
  CSD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return CSD_OK;
}
int csd_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return CSD_OK;
  CSD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return CSD_OK;
}
int csd_memset(void *dst, int value, size_t bytes, void *stream) {
  if (bytes == 0) return CSD_OK;
  CSD_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  return CSD_OK;
}
int csd_stream_synchronize(void *stream) {
  CSD_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CSD_OK;
}
int csd_stream_create(void **stream) {
  if (!stream) return fail(CSD_ERR_INVALID_ARGUMENT, "null pointer");
  hipStream_t s;
  CSD_HIP(hipStreamCreate(&s));
  *stream = (void *)s;
  return CSD_OK;
}
int csd_stream_destroy(void *stream) {
  if (stream) CSD_HIP(hipStreamDestroy((hipStream_t)stream));
  return CSD_OK;
}

int csd_init(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return fail(CSD_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count)
    return fail(CSD_ERR_INVALID_ARGUMENT, "device index out of range");
  CSD_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  CSD_HIP(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(CSD_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                       ", this library is built for gfx950");
  return CSD_OK;
}

int csd_plant_create(const csd_plant_desc *d, csd_plant **out) {
  if (!d || !out) return fail(CSD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const int L = d->num_links, J = d->num_joints, G = d->num_geometries,
            P = d->num_collision_pairs, dof = d->num_configuration_variables;
  if (L < 1 || J < 0 || G < 0 || P < 0 || dof < 0)
    return fail(CSD_ERR_INVALID_ARGUMENT, "negative or empty plant sizes");
  if (dof > kMaxDof)
    return fail(CSD_ERR_CAPACITY, "more than 32 configuration variables");
  if (J > 0 && (!d->joint_type || !d->joint_parent_link || !d->joint_child_link ||
                !d->joint_config_idx || !d->traversal || !d->joint_X_PL_J ||
                !d->joint_axis))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null joint array");
  if (G > 0 && (!d->geom_type || !d->geom_link || !d->geom_dims || !d->geom_X_L_B))
    return fail(CSD_ERR_INVALID_ARGUMENT, "null geometry array");
  if (P > 0 && !d->collision_pairs)
    return fail(CSD_ERR_INVALID_ARGUMENT, "null pair array");

  // ---- fold fixed joints: per link (body, pose in body frame) ----
  std::vector<int> link_body(L, -2);
  std::vector<Aff> link_off(L, Aff::identity());
  link_body[0] = -1;  // root is the world frame
  std::vector<DevBody> bodies;
  std::vector<int> seen(J, 0);
  for (int t = 0; t < J; ++t) {
    int j = d->traversal[t];
    if (j < 0 || j >= J || seen[j])
      return fail(CSD_ERR_INVALID_ARGUMENT, "traversal is not a permutation of joints");
    seen[j] = 1;
    int pl = d->joint_parent_link[j], cl = d->joint_child_link[j];
    if (pl < 0 || pl >= L || cl < 0 || cl >= L)
      return fail(CSD_ERR_INVALID_ARGUMENT, "joint link index out of range");
    if (link_body[pl] == -2)
      return fail(CSD_ERR_INVALID_ARGUMENT, "traversal visits a joint before its parent link");
    Aff XPJ = Aff::from_colmajor4(d->joint_X_PL_J + 16 * j);
    Aff C = link_off[pl] * XPJ;
    int type = d->joint_type[j];
    if (type == CSD_JOINT_FIXED) {
      link_body[cl] = link_body[pl];
      link_off[cl] = C;
    } else if (type == CSD_JOINT_REVOLUTE || type == CSD_JOINT_PRISMATIC) {
      int qi = d->joint_config_idx[j];
      if (qi < 0 || qi >= dof)
        return fail(CSD_ERR_INVALID_ARGUMENT, "joint configuration index out of range");
      if ((int)bodies.size() >= kMaxBodies)
        return fail(CSD_ERR_CAPACITY, "more than 64 moving joints");
      DevBody b;
      memset(&b, 0, sizeof(b));
      b.parent = link_body[pl];
      b.type = type;
      b.q_idx = qi;
      C.to_float(b.C);
      const float *af = d->joint_axis + 3 * j;
      double a[3] = {af[0], af[1], af[2]};
      if (type == CSD_JOINT_REVOLUTE) {
        double K[9] = {0, -a[2], a[1], a[2], 0, -a[0], -a[1], a[0], 0};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double ck = 0, ca = 0;
            for (int k = 0; k < 3; ++k) {
              ck += C.m[r * 4 + k] * K[k * 3 + c];
              ca += C.m[r * 4 + k] * a[k] * a[c];
            }
            b.CK[r * 3 + c] = (float)ck;
            b.CAm[r * 3 + c] = (float)(ca - C.m[r * 4 + c]);
          }
      } else {
        for (int r = 0; r < 3; ++r)
          b.CK[r] = (float)(C.m[r * 4 + 0] * a[0] + C.m[r * 4 + 1] * a[1] + C.m[r * 4 + 2] * a[2]);
      }
      link_body[cl] = (int)bodies.size();
      link_off[cl] = Aff::identity();
      bodies.push_back(b);
    } else {
      return fail(CSD_ERR_INVALID_ARGUMENT, "unknown joint type");
    }
  }
  for (int l = 0; l < L; ++l)
    if (link_body[l] == -2)
      return fail(CSD_ERR_INVALID_ARGUMENT, "link not reachable from the root");
  const int NB = (int)bodies.size();

  // ---- geometries ----
  std::vector<DevGeom> geoms(G);
  std::vector<Aff> gpose(G);
  std::vector<int> gbody(G);
  std::vector<std::array<double, 3>> gdim(G);
  for (int g = 0; g < G; ++g) {
    int l = d->geom_link[g], t = d->geom_type[g];
    if (l < 0 || l >= L) return fail(CSD_ERR_INVALID_ARGUMENT, "geometry link out of range");
    if (t != CSD_BOX && t != CSD_SPHERE && t != CSD_CAPSULE)
      return fail(CSD_ERR_UNSUPPORTED, "cylinders are not supported (box, sphere, capsule are)");
    const float *dm = d->geom_dims + 3 * g;
    DevGeom &dg = geoms[g];
    memset(&dg, 0, sizeof(dg));
    dg.type = t;
    dg.body = link_body[l];
    gbody[g] = dg.body;
    gpose[g] = link_off[l] * Aff::from_colmajor4(d->geom_X_L_B + 16 * g);
    gpose[g].to_float(dg.G);
    if (t == CSD_BOX) {
      for (int k = 0; k < 3; ++k) {
        dg.dims[k] = dm[k] * 0.5f;
        gdim[g][k] = (double)dm[k] * 0.5;
      }
    } else if (t == CSD_SPHERE) {
      dg.dims[0] = dm[0];
      gdim[g] = {dm[0], 0, 0};
    } else {
      dg.dims[0] = dm[0];
      dg.dims[1] = dm[1] * 0.5f;
      gdim[g] = {dm[0], (double)dm[1] * 0.5, 0};
    }
  }

  // ---- pairs ----
  std::vector<int32_t> pairs, pair_types;
  int always_colliding = 0;
  for (int p = 0; p < P; ++p) {
    int a = d->collision_pairs[2 * p], b = d->collision_pairs[2 * p + 1];
    if (a < 0 || a >= G || b < 0 || b >= G)
      return fail(CSD_ERR_INVALID_ARGUMENT, "collision pair index out of range");
    if (gbody[a] == gbody[b]) {
      // relative pose is constant: decide once
      if (static_pair_sd(geoms[a].type, gpose[a], gdim[a].data(), geoms[b].type, gpose[b],
                         gdim[b].data()) < 0)
        always_colliding = 1;
      continue;
    }
    if (type_rank(geoms[a].type) > type_rank(geoms[b].type)) std::swap(a, b);
    pairs.push_back(a);
    pairs.push_back(b);
    pair_types.push_back(type_rank(geoms[a].type) * 4 + type_rank(geoms[b].type));
  }
  const int NP = (int)pair_types.size();

  // ---- fast-path tables: every moving geometry in a pair is a sphere ----
  // the fast kernel also needs chain-ordered bodies (parent is the previous
  // body or the world), so FK keeps only the current frame live
  bool fast = NB <= 16;
  for (int b = 0; b < NB && fast; ++b)
    if (bodies[b].parent != -1 && bodies[b].parent != b - 1) fast = false;
  std::vector<int> per_body(NB, 0);
  // moving pair geometry must be spheres of radius below 1 m (the fast
  // sphere-box test clamps axis distances to [0, 1])
  for (int p = 0; p < NP && fast; ++p)
    for (int s = 0; s < 2; ++s) {
      int g = pairs[2 * p + s];
      if (gbody[g] >= 0 && (geoms[g].type != CSD_SPHERE || !(geoms[g].dims[0] < 1.f)))
        fast = false;
    }
  for (int g = 0; g < G; ++g)
    if (gbody[g] >= 0 && geoms[g].type == CSD_SPHERE) per_body[gbody[g]]++;
  int maxk = 0;
  for (int c : per_body) maxk = std::max(maxk, c);
  int kmax = maxk <= 4 ? 4 : (maxk <= 8 ? 8 : 0);
  if (kmax == 0) fast = false;
  int nbt = NB <= 8 ? 8 : 16;
  if (fast && nbt * kmax > kFastMaxSlots) fast = false;
  if (fast && dof > nbt) fast = false;  // the fast kernels hold q in nbt registers
  const int nslots = fast ? nbt * kmax : 0;
  std::vector<float4> slot_sphere(std::max(nslots, 1), make_float4(0, 0, 0, -1));
  std::vector<int32_t> geom_slot(std::max(G, 1), -1);
  std::vector<DevObstacle> obstacles;
  std::vector<float4> obstacle_bound;
  std::vector<int32_t> obstacle_type;
  std::vector<std::pair<int, int>> obstacle_slot;  // (obstacle, slot) pairs
  const int nbt_alloc = fast ? nbt : 1;
  std::vector<uint64_t> self_active(nbt_alloc, 0);
  std::vector<uint64_t> self_bits((size_t)nbt_alloc * nbt_alloc, 0);
  const bool tables = fast && nslots <= 64;
  std::vector<float> self_r2(tables ? (size_t)nbt * nbt * kmax * kmax : 1, 0.f);
  int num_self_pairs = 0;
  // body bounds for culling: per body, the sphere whose centre encloses all
  // of the body's spheres in the smallest radius takes the body's first slot
  // and body_bound[b] holds that radius, inflated; the kernels then skip a
  // body's obstacle tests when, for every undecided lane, its first sphere
  // centre is farther from the obstacle's bounding sphere than the sum.
  std::vector<float> body_bound(nbt_alloc, 0.f);
  if (fast) {
    std::vector<int> used(NB, 0), first(NB, -1);
    std::vector<double> best(NB, 0);
    for (int g = 0; g < G; ++g) {
      if (gbody[g] < 0 || geoms[g].type != CSD_SPHERE) continue;
      const int b = gbody[g];
      double rad = 0;
      for (int h = 0; h < G; ++h) {
        if (gbody[h] != b || geoms[h].type != CSD_SPHERE) continue;
        double d2 = 0;
        for (int k = 0; k < 3; ++k) {
          const double dk = (double)geoms[h].G[4 * k + 3] - geoms[g].G[4 * k + 3];
          d2 += dk * dk;
        }
        rad = std::max(rad, std::sqrt(d2) + geoms[h].dims[0]);
      }
      if (first[b] < 0 || rad < best[b]) {
        first[b] = g;
        best[b] = rad;
      }
    }
    for (int b = 0; b < NB; ++b) {
      used[b] = first[b] >= 0 ? 1 : 0;
      body_bound[b] = (float)(best[b] * (1.0 + kBoundInflate) + kBoundPad);
    }
    for (int g = 0; g < G; ++g) {
      if (gbody[g] < 0 || geoms[g].type != CSD_SPHERE) continue;
      int b = gbody[g];
      int slot = b * kmax + (g == first[b] ? 0 : used[b]++);
      geom_slot[g] = slot;
      slot_sphere[slot] = make_float4(geoms[g].G[3], geoms[g].G[7], geoms[g].G[11], geoms[g].dims[0]);
    }
    std::vector<int> obs_of_geom(G, -1);
    for (int p = 0; p < NP; ++p) {
      int a = pairs[2 * p], b = pairs[2 * p + 1];
      int ma = gbody[a] >= 0, mb = gbody[b] >= 0;
      if (ma && mb) {
        int sa = geom_slot[a], sb = geom_slot[b];
        if (sa > sb) std::swap(sa, sb);
        const int ba = sa / kmax, bb = sb / kmax, ka = sa % kmax, kb = sb % kmax;
        self_active[ba] |= (uint64_t)1 << bb;
        self_bits[(size_t)ba * nbt + bb] |= (uint64_t)1 << (ka * kmax + kb);
        if (tables) {
          const float rs = geoms[a].dims[0] + geoms[b].dims[0];
          self_r2[(((size_t)ba * nbt + bb) * kmax + ka) * kmax + kb] = rs * rs;
        }
        ++num_self_pairs;
        continue;
      }
      int sg = ma ? a : b;  // moving sphere
      int og = ma ? b : a;  // world-fixed obstacle
      if (obs_of_geom[og] < 0) {
        obs_of_geom[og] = (int)obstacle_type.size();
        DevObstacle rec;
        memset(&rec, 0, sizeof(rec));
        const Aff &X = gpose[og];
        int t = geoms[og].type;
        if (t == CSD_BOX) {
          // inverse pose: local = R^T c - R^T p
          for (int r = 0; r < 3; ++r) {
            double tr = 0;
            for (int c = 0; c < 3; ++c) {
              rec.v[r * 3 + c] = (float)X.m[c * 4 + r];
              tr -= X.m[c * 4 + r] * X.m[c * 4 + 3];
            }
            rec.v[9 + r] = (float)tr;
            rec.v[12 + r] = (float)gdim[og][r];
          }
        } else if (t == CSD_SPHERE) {
          rec.v[0] = (float)X.m[3];
          rec.v[1] = (float)X.m[7];
          rec.v[2] = (float)X.m[11];
          rec.v[3] = (float)gdim[og][0];
        } else {  // capsule: p0, d = p1 - p0, 1/|d|^2, radius
          double hl = gdim[og][1];
          double p0[3], dd[3], n2 = 0;
          for (int k = 0; k < 3; ++k) {
            p0[k] = X.m[k * 4 + 3] - hl * X.m[k * 4 + 2];
            dd[k] = 2 * hl * X.m[k * 4 + 2];
            n2 += dd[k] * dd[k];
          }
          for (int k = 0; k < 3; ++k) {
            rec.v[k] = (float)p0[k];
            rec.v[3 + k] = (float)dd[k];
          }
          rec.v[6] = n2 > 0 ? (float)(1.0 / n2) : 0.f;
          rec.v[7] = (float)gdim[og][0];
        }
        // bounding sphere: box centre and |half extents|, the sphere itself,
        // capsule midpoint and half length + radius; inflated like the bodies'
        double bc[3] = {X.m[3], X.m[7], X.m[11]}, br;
        if (t == CSD_BOX)
          br = std::sqrt(gdim[og][0] * gdim[og][0] + gdim[og][1] * gdim[og][1] +
                         gdim[og][2] * gdim[og][2]);
        else if (t == CSD_SPHERE)
          br = gdim[og][0];
        else
          br = gdim[og][0] + gdim[og][1];
        obstacle_bound.push_back(make_float4((float)bc[0], (float)bc[1], (float)bc[2],
                                             (float)(br * (1.0 + kBoundInflate) + kBoundPad)));
        obstacles.push_back(rec);
        obstacle_type.push_back(t);
      }
      obstacle_slot.emplace_back(obs_of_geom[og], geom_slot[sg]);
    }
  }
  std::vector<uint64_t> obstacle_mask(std::max<size_t>(2 * obstacles.size(), 2), 0);
  for (const auto &os : obstacle_slot)
    obstacle_mask[2 * os.first + (os.second >> 6)] |= (uint64_t)1 << (os.second & 63);
  // contact distances in FP32, as the device expressions r * r and
  // (r + r_o) * (r + r_o) round them
  std::vector<float> obstacle_r2(tables ? std::max<size_t>(obstacles.size() * nslots, 1) : 1, -1.f);
  if (tables)
    for (const auto &os : obstacle_slot) {
      const int o = os.first, sl = os.second;
      const float r = slot_sphere[sl].w;
      float r2;
      if (obstacle_type[o] == CSD_BOX) {
        r2 = r * r;
      } else {
        const float rs = r + obstacles[o].v[obstacle_type[o] == CSD_SPHERE ? 3 : 7];
        r2 = rs * rs;
      }
      obstacle_r2[(size_t)o * nslots + sl] = r2;
    }

  // ---- links for standalone FK ----
  std::vector<DevLink> links(L);
  for (int l = 0; l < L; ++l) {
    memset(&links[l], 0, sizeof(DevLink));
    links[l].body = link_body[l];
    link_off[l].to_float(links[l].O);
  }

  // ---- upload ----
  Builder bld;
  size_t o_bodies = bld.add(bodies.data(), bodies.size());
  size_t o_geoms = bld.add(geoms.data(), geoms.size());
  size_t o_pairs = bld.add(pairs.data(), pairs.size());
  size_t o_ptypes = bld.add(pair_types.data(), pair_types.size());
  size_t o_links = bld.add(links.data(), links.size());
  size_t o_slots = bld.add(slot_sphere.data(), slot_sphere.size());
  size_t o_obs = bld.add(obstacles.data(), obstacles.size());
  size_t o_otype = bld.add(obstacle_type.data(), obstacle_type.size());
  size_t o_omask = bld.add(obstacle_mask.data(), obstacle_mask.size());
  size_t o_sact = bld.add(self_active.data(), self_active.size());
  size_t o_sbits = bld.add(self_bits.data(), self_bits.size());
  size_t o_sr2 = bld.add(self_r2.data(), self_r2.size());
  size_t o_or2 = bld.add(obstacle_r2.data(), obstacle_r2.size());
  size_t o_gslot = bld.add(geom_slot.data(), geom_slot.size());
  if (obstacle_bound.empty()) obstacle_bound.push_back(make_float4(0, 0, 0, 0));
  size_t o_obound = bld.add(obstacle_bound.data(), obstacle_bound.size());
  size_t o_bbound = bld.add(body_bound.data(), body_bound.size());

  csd_plant *pl = new csd_plant();
  hipError_t e = hipGetDevice(&pl->device);
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "hipGetDevice");
  }
  e = hipMalloc(&pl->arena, bld.bytes.size());
  if (e != hipSuccess) {
    delete pl;
    return hip_fail(e, "hipMalloc(plant image)");
  }
  e = hipMemcpy(pl->arena, bld.bytes.data(), bld.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(pl->arena);
    delete pl;
    return hip_fail(e, "hipMemcpy(plant image)");
  }
  char *base = (char *)pl->arena;
  PlantImage &img = pl->img;
  memset(&img, 0, sizeof(img));
  img.num_links = L;
  img.dof = dof;
  img.num_bodies = NB;
  img.num_geoms = G;
  img.num_pairs = NP;
  img.always_colliding = always_colliding;
  img.fast = fast ? 1 : 0;
  img.kmax = fast ? kmax : 0;
  img.nslots = nslots;
  img.num_obstacles = (int)obstacle_type.size();
  img.num_self_pairs = num_self_pairs;
  img.q_identity = 1;
  for (int bi = 0; bi < NB; ++bi)
    if (bodies[bi].q_idx != bi) img.q_identity = 0;
  img.bodies = (const DevBody *)(base + o_bodies);
  img.geoms = (const DevGeom *)(base + o_geoms);
  img.pairs = (const int32_t *)(base + o_pairs);
  img.pair_types = (const int32_t *)(base + o_ptypes);
  img.links = (const DevLink *)(base + o_links);
  img.slot_sphere = (const float4 *)(base + o_slots);
  img.obstacles = (const DevObstacle *)(base + o_obs);
  img.obstacle_type = (const int32_t *)(base + o_otype);
  img.obstacle_mask = (const uint64_t *)(base + o_omask);
  img.self_active = (const uint64_t *)(base + o_sact);
  img.self_bits = (const uint64_t *)(base + o_sbits);
  img.self_r2 = (const float *)(base + o_sr2);
  img.obstacle_r2 = (const float *)(base + o_or2);
  img.geom_slot = (const int32_t *)(base + o_gslot);
  img.obstacle_bound = (const float4 *)(base + o_obound);
  img.body_bound = (const float *)(base + o_bbound);
  pl->geom_type.assign(d->geom_type, d->geom_type + G);
  pl->geom_link.assign(d->geom_link, d->geom_link + G);
  pl->geom_body = gbody;
  pl->geom_slot.assign(geom_slot.begin(), geom_slot.begin() + G);
  pl->num_configuration_variables = dof;
  pl->num_links = L;
  *out = pl;
  return CSD_OK;
}

int csd_plant_destroy(csd_plant *plant) {
  if (!plant) return CSD_OK;
  // kernels launched asynchronously may still read the image
  hipDeviceSynchronize();
  if (plant->arena) hipFree(plant->arena);
  if (plant->screen_arena) hipFree(plant->screen_arena);
  delete plant;
  return CSD_OK;
}

int csd_plant_info(const csd_plant *plant, int32_t *num_links, int32_t *dof,
                   int32_t *num_pairs, int32_t *fast_path) {
  if (!plant) return fail(CSD_ERR_INVALID_ARGUMENT, "null plant");
  if (num_links) *num_links = plant->img.num_links;
  if (dof) *dof = plant->img.dof;
  if (num_pairs) *num_pairs = plant->img.num_pairs;
  if (fast_path) *fast_path = plant->img.fast;
  return CSD_OK;
}

}  // extern "C"
