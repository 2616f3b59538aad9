// swd_fmst.hip -- FMST traveltimes and ray Fréchet weights on CDNA4: the GPU
// form of the per-station-pair sequence fmst_init / fmst_run / fmst_reset /
// fmst_raypath that SurfTime::travel_time, frechet_matrix and compute_grad
// run in their OpenMP loops (src/SWD/SurfaceWave.cpp:518-540, 612-624,
// 737-760; src/SWD/fmst.f90, src/SWD/fmm.f90).
//
// The method is the reference's, in single precision and the reference's
// operation order (no FMA contraction in this file):
//  * the dispersion map (B-spline control values, column n = ix + iy*nx) is
//    diced 8x8 into the propagation grid by cubic B-splines (gridder);
//  * around the source a box of +-8 propagation nodes is diced 8x8 again
//    (bsplrefine) and a fast-marching pass runs there until the front
//    reaches the box edge (travel, urg = 1);
//  * the box's nodes seed the narrow band of a second fast-marching pass over
//    the whole grid (urg = 2), with the mixed-order upwind update (fouds2)
//    and a binary-heap band (addtree/downtree/updtree);
//  * receiver times by bilinear interpolation (srtimes);
//  * rays from each receiver down the traveltime gradient in steps of half a
//    cell, with the cubic B-spline weights -int B/v^2 dl of the group map
//    accumulated per segment, and the group time int dl/v (fmst_raypath).
//
// Fast marching is a sequence: each accepted node decides the next.  So a
// station pair is one workgroup of one wavefront whose first lane runs the
// march, with the heap's first 2048 entries in LDS (the band of a few
// hundred nodes lives there; deeper entries spill to global memory up to the
// reference's own capacity), and each accepted node's whole update stencil
// -- the 25 nodes within three steps -- read in one batch of independent
// loads.  The lanes share the parallel parts (refined velocities, band
// seeding, receiver times).  Thousands of pairs run at once, one wavefront
// each.  Rays are one lane each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/jdsg_swd.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;
int fail(const std::string &m) {
  g_err = m;
  return 1;
}
#define FCHK(x)                                                                  \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) return fail(std::string("hip: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr float kPi = 3.1415926535898f;  // globalp: REAL(sp), PARAMETER :: pi
constexpr int kHeapLds = 1536;            // heap entries held in LDS per pair
constexpr int kRmax = 129;                // refined nodes per axis: 16 cells x 8 + 1

struct __attribute__((aligned(8))) Node {
  float t;  // traveltime
  int s;    // -1 far, 0 alive, > 0 position in the heap
};
struct __attribute__((aligned(8))) HEnt {
  float v;  // traveltime of the node (always equal to its Node::t)
  int n;    // node, (ix << 16) | iz, 1-based
};

// fmst_init constants (fmst.f90:9-82), coarse propagation grid
struct FGrid {
  int nx, ny, nvx, nvz, nnx, nnz, gdx, gdz, sgdl;
  float gox, goz, dnx, dnz, dvx, dvz, earth, snb;
  float dpl_sr;  // srtimes' source neighbourhood (fmm.f90:1027-1031)
  float dpl;     // fmst_raypath's step, half of that (fmst.f90:459-464)
};

// one station pair, constants of fmst_run computed on the host in the
// reference's float order (fmst.f90:140-197)
struct FPair {
  int idx, plane, r0, nr;
  float srcx, srcz;
  int vnl, vnr, vnt, vnb, nrnx, nrnz;
  float gorx, gorz, drnx, drnz;
  int isxr, iszr;  // source cell of the refined march (travel, adjusted)
  int isxq, iszq;  // source cell on the refined grid as fmst_raypath takes it
  int maxbt;       // heap capacity
  int status;      // 1: source outside the model
  int rayt;        // the receivers' times are the ray times (Rg, Lg)
};

__device__ __forceinline__ void basis4(float u, float b[4]) {
  const float u2 = u * u, u3 = u * u2, w = 1.0f - u;
  b[0] = (w * (w * w)) / 6.0f;
  b[1] = (4.0f - 6.0f * u2 + 3.0f * u3) / 6.0f;
  b[2] = (1.0f + 3.0f * u + 3.0f * u2 - 3.0f * u3) / 6.0f;
  b[3] = u3 / 6.0f;
}

// ---------------------------------------------------------------- gridder
// gridder (fmm.f90:776-870) of map column `idx` into plane p: one thread per
// propagation node; node (stz, stx) lies in B-spline cell (i, j) at local
// (l, m), the last node of an axis in the last cell at l = gd + 1.
__global__ void __launch_bounds__(256)
fmst_gridder(FGrid g, int nplane, const int *plane_idx, const double *map, long long kstride,
             float *out) {
  const int nn = g.nnx * g.nnz;
  const int p = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nplane || q >= nn) return;
  const double *pv = map + (long long)plane_idx[p] * kstride;
  const int stz = q % g.nnz + 1, stx = q / g.nnz + 1;
  const int i = min((stz - 1) / g.gdz + 1, g.nvz - 1), l = stz - g.gdz * (i - 1);
  const int j = min((stx - 1) / g.gdx + 1, g.nvx - 1), m = stx - g.gdx * (j - 1);
  float ui[4], vi[4];
  basis4((float)(m - 1) / (float)g.gdx, ui);
  basis4((float)(l - 1) / (float)g.gdz, vi);
  float sumi = 0.0f;
#pragma unroll
  for (int i1 = 1; i1 <= 4; ++i1) {
    float sumj = 0.0f;
#pragma unroll
    for (int j1 = 1; j1 <= 4; ++j1)
      sumj = sumj + ui[j1 - 1] * (float)pv[(i - 2 + i1) * (g.nvx + 2) + (j - 2 + j1)];
    sumi = sumi + vi[i1 - 1] * sumj;
  }
  out[(long long)p * nn + q] = sumi;
}

// ---------------------------------------------------------------- march
// The heap band.  LdsHeap holds it in LDS (the fast path: the band of a
// 241 x 265 grid stays within a few hundred nodes); a pair whose band
// outgrows that is marched again with GlobalHeap, in global memory up to
// the reference's own capacity.
typedef __attribute__((address_space(3))) HEnt LdsHEnt;  // LDS-typed: ds_* access
struct LdsHeap {
  LdsHEnt *e;  // entries 1..cap
  int ntr, cap;
  __device__ __forceinline__ HEnt get(int k) const {
    HEnt r;
    r.v = e[k - 1].v;
    r.n = e[k - 1].n;
    return r;
  }
  __device__ __forceinline__ void put(int k, HEnt v) {
    e[k - 1].v = v.v;
    e[k - 1].n = v.n;
  }
};
struct GlobalHeap {
  HEnt *e;
  int ntr, cap;
  __device__ __forceinline__ HEnt get(int k) const { return e[k - 1]; }
  __device__ __forceinline__ void put(int k, HEnt v) { e[k - 1] = v; }
};

// a grid being marched: nodes (iz, ix) at nd[(iz-1) + (ix-1)*nnz]
struct MGrid {
  Node *nd;
  const float *vel;    // same layout
  const float *risti;  // earth * sin(gox + (ix-1)*dnx), per column
  int nnx, nnz;
  float gox, goz, dnx, dnz, earth;
};

__device__ __forceinline__ long long nid(const MGrid &G, int iz, int ix) {
  return (long long)(iz - 1) + (long long)(ix - 1) * G.nnz;
}
__device__ __forceinline__ int pack(int iz, int ix) { return (ix << 16) | iz; }

// The march runs on the whole wavefront with wave-uniform control: heap
// entries read from LDS are made uniform (readfirstlane), so the heap's
// sifts branch on scalars; every lane performs the same stores, so each
// lane's later loads follow its own stores in program order.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
template <class Heap>
__device__ __forceinline__ HEnt hget(const Heap &H, int k) {
  HEnt e = H.get(k);
  e.v = uni(e.v);
  e.n = uni(e.n);
  return e;
}

// the up to four neighbours whose heap positions a sift must keep current
struct Pending {
  int n[4], pos[4];
  __device__ __forceinline__ void moved(int node, int k) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (n[i] == node) pos[i] = k;
  }
};

__device__ __forceinline__ void set_pos(const MGrid &G, int node, int k) {
  G.nd[nid(G, node & 0xffff, node >> 16)].s = k;
}

// addtree (fmm.f90:616-653): the node's traveltime t is already in place
template <class Heap>
__device__ bool heap_add(Heap &H, const MGrid &G, int node, float t, Pending &pd) {
  if (H.ntr + 1 > H.cap) return false;
  H.ntr += 1;
  int k = H.ntr;
  while (k > 1) {
    const int p = k / 2;
    const HEnt e = hget(H, p);
    if (t < e.v) {
      H.put(k, e);
      set_pos(G, e.n, k);
      pd.moved(e.n, k);
      k = p;
    } else {
      break;
    }
  }
  H.put(k, HEnt{t, node});
  set_pos(G, node, k);
  pd.moved(node, k);
  return true;
}

// updtree (fmm.f90:745-772) from position k, new traveltime t
template <class Heap>
__device__ void heap_up(Heap &H, const MGrid &G, int node, float t, int k, Pending &pd) {
  while (k > 1) {
    const int p = k / 2;
    const HEnt e = hget(H, p);
    if (t < e.v) {
      H.put(k, e);
      set_pos(G, e.n, k);
      pd.moved(e.n, k);
      k = p;
    } else {
      break;
    }
  }
  H.put(k, HEnt{t, node});
  set_pos(G, node, k);
  pd.moved(node, k);
}

// downtree (fmm.f90:662-738) after the root was taken
template <class Heap>
__device__ void heap_pop(Heap &H, const MGrid &G) {
  if (H.ntr == 1) {
    H.ntr = 0;
    return;
  }
  const HEnt x = hget(H, H.ntr);
  H.ntr -= 1;
  const int ntr = H.ntr;
  int hole = 1, c = 2;
  while (c < ntr) {
    HEnt a = hget(H, c);
    const HEnt b = hget(H, c + 1);
    if (a.v > b.v) {
      c = c + 1;
      a = b;
    }
    if (a.v < x.v) {
      H.put(hole, a);
      set_pos(G, a.n, hole);
      hole = c;
      c = 2 * hole;
    } else {
      c = ntr + 1;
    }
  }
  if (c == ntr) {
    const HEnt a = hget(H, c);
    if (a.v < x.v) {
      H.put(hole, a);
      set_pos(G, a.n, hole);
      hole = c;
    }
  }
  H.put(hole, x);
  set_pos(G, x.n, hole);
}

// One quadrant (j = ix + sj, k = iz + sk) of fouds2 (fmm.f90:437-608) at
// node (iz, ix): quad_load reads its nodes (bounds-checked, before the
// node's own status is known, so the loads of a step go out together);
// quad_eval gives the trial time of its solution, or +inf where the
// quadrant gives none (out of the grid, or no alive upwind node).  The
// reference's minimum over the quadrants does not depend on their order.
struct QLoad {
  Node nj, nj2, nk, nk2;
  float vel, risti;
  bool jin, kin, j2in, k2in;
};

__device__ __forceinline__ QLoad quad_load(const MGrid &G, int iz, int ix, int sj, int sk, bool in) {
  QLoad L;
  const int j = ix + sj, k = iz + sk, j2 = j + sj, k2 = k + sk;
  L.jin = in && j >= 1 && j <= G.nnx;
  L.kin = in && k >= 1 && k <= G.nnz;
  L.j2in = L.jin && (sj < 0 ? (j2 >= 1) : (j2 <= G.nnx));
  L.k2in = L.kin && (sk < 0 ? (k2 >= 1) : (k2 <= G.nnz));
  const Node none{0.0f, -2};
  L.nj = L.jin ? G.nd[nid(G, iz, j)] : none;
  L.nk = L.kin ? G.nd[nid(G, k, ix)] : none;
  L.nj2 = L.j2in ? G.nd[nid(G, iz, j2)] : none;
  L.nk2 = L.k2in ? G.nd[nid(G, k2, ix)] : none;
  L.vel = in ? G.vel[nid(G, iz, ix)] : 1.0f;
  L.risti = in ? G.risti[ix - 1] : 0.0f;
  return L;
}

__device__ __forceinline__ float quad_eval(const MGrid &G, const QLoad &L) {
  const float kInf = __int_as_float(0x7f800000);
  if (!L.jin || !L.kin) return kInf;
  const Node nj = L.nj, nj2 = L.nj2, nk = L.nk, nk2 = L.nk2;
  const float risti = L.risti;
  const float slown = 1.0f / L.vel;
  const float ri = G.earth;
  const int swj = (nj.s == 0 && L.j2in && nj2.s == 0 && nj.t > nj2.t) ? 0 : -1;
  const int swk = (nk.s == 0 && L.k2in && nk2.s == 0 && nk.t > nk2.t) ? 0 : -1;
  float a, b, c, u, v, em, tref, tdiv;
  if (swj == 0) {
    if (swk == 0) {
      u = 2.0f * ri * G.dnx;
      v = 2.0f * risti * G.dnz;
      em = 4.0f * nj.t - nj2.t - 4.0f * nk.t;
      em = em + nk2.t;
      a = v * v + u * u;
      b = 2.0f * em * (u * u);
      c = (u * u) * (em * em - (slown * slown) * (v * v));
      tref = 4.0f * nj.t - nj2.t;
      tdiv = 3.0f;
    } else if (nk.s == 0) {
      u = risti * G.dnz;
      v = 2.0f * ri * G.dnx;
      em = 3.0f * nk.t - 4.0f * nj.t + nj2.t;
      a = v * v + 9.0f * (u * u);
      b = 6.0f * em * (u * u);
      c = (u * u) * (em * em - (slown * slown) * (v * v));
      tref = nk.t;
      tdiv = 1.0f;
    } else {
      u = 2.0f * ri * G.dnx;
      a = 1.0f;
      b = 0.0f;
      c = -((u * u) * (slown * slown));
      tref = 4.0f * nj.t - nj2.t;
      tdiv = 3.0f;
    }
  } else if (nj.s == 0) {
    if (swk == 0) {
      u = ri * G.dnx;
      v = 2.0f * risti * G.dnz;
      em = 3.0f * nj.t - 4.0f * nk.t + nk2.t;
      a = v * v + 9.0f * (u * u);
      b = 6.0f * em * (u * u);
      c = (u * u) * (em * em - (v * v) * (slown * slown));
      tref = nj.t;
      tdiv = 1.0f;
    } else if (nk.s == 0) {
      u = ri * G.dnx;
      v = risti * G.dnz;
      em = nk.t - nj.t;
      a = u * u + v * v;
      b = -(2.0f * (u * u) * em);
      c = (u * u) * (em * em - (v * v) * (slown * slown));
      tref = nj.t;
      tdiv = 1.0f;
    } else {
      a = 1.0f;
      b = 0.0f;
      c = -((slown * slown) * (ri * ri) * (G.dnx * G.dnx));
      tref = nj.t;
      tdiv = 1.0f;
    }
  } else if (swk == 0) {
    u = 2.0f * risti * G.dnz;
    a = 1.0f;
    b = 0.0f;
    c = -((u * u) * (slown * slown));
    tref = 4.0f * nk.t - nk2.t;
    tdiv = 3.0f;
  } else if (nk.s == 0) {
    a = 1.0f;
    b = 0.0f;
    c = -((slown * slown) * (risti * risti) * (G.dnz * G.dnz));
    tref = nk.t;
    tdiv = 1.0f;
  } else {
    return kInf;
  }
  float rd1 = b * b - 4.0f * a * c;
  if (rd1 < 0.0f) rd1 = 0.0f;
  const float tdsh = (-b + sqrtf(rd1)) / (2.0f * a);
  return (tref + tdsh) / tdiv;
}

// the fast-marching loop of travel (fmm.f90:210-328); urg = 1 adds the
// refined grid's exit test (:216-237, comparing the box bounds with the
// refined node counts, as written).  Each accepted node's four neighbours
// take their trial times from 16 lanes, one quadrant each, reduced to the
// minimum per neighbour; the heap updates follow in the reference's order
// (x - 1, x + 1, z - 1, z + 1).  Returns false on heap overflow.
template <class Heap>
__device__ bool march(Heap &H, const MGrid &G, int urg, int vnl, int vnr, int vnt, int vnb) {
  const int lane = threadIdx.x & 15;
  const int q = lane >> 2;                                  // this lane's neighbour
  const int sj = (lane & 1) ? 1 : -1, sk = (lane & 2) ? 1 : -1;  // and quadrant
  const int qdz = (q == 2) ? -1 : (q == 3) ? 1 : 0, qdx = (q == 0) ? -1 : (q == 1) ? 1 : 0;
  while (H.ntr > 0) {
    const HEnt r = hget(H, 1);
    const int ix = r.n >> 16, iz = r.n & 0xffff;
    if (urg == 1) {
      int swrg = 0;
      if (ix == 1 && vnl != 1) swrg = 1;
      if (ix == G.nnx && vnr != G.nnx) swrg = 1;
      if (iz == 1 && vnt != 1) swrg = 1;
      if (iz == G.nnz && vnb != G.nnz) swrg = 1;
      if (swrg == 1) {
        G.nd[nid(G, iz, ix)].s = 0;
        return true;
      }
    }
    G.nd[nid(G, iz, ix)].s = 0;
    heap_pop(H, G);
    // each lane: its neighbour's state and one quadrant's trial time
    const int nz = iz + qdz, nx = ix + qdx;
    const bool in = nz >= 1 && nz <= G.nnz && nx >= 1 && nx <= G.nnx;
    const int st = in ? G.nd[nid(G, nz, nx)].s : -2;
    const QLoad ql = quad_load(G, nz, nx, sj, sk, in);
    float cand = __int_as_float(0x7f800000);
    if (in && st != 0) cand = quad_eval(G, ql);
    {
      float o = __shfl_xor(cand, 1);
      cand = (o < cand) ? o : cand;
      o = __shfl_xor(cand, 2);
      cand = (o < cand) ? o : cand;
    }
    Pending pd;
    float tq[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      tq[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cand), 4 * i));
      pd.pos[i] = __builtin_amdgcn_readlane(st, 4 * i);
    }
    pd.n[0] = pack(iz, ix - 1);
    pd.n[1] = pack(iz, ix + 1);
    pd.n[2] = pack(iz - 1, ix);
    pd.n[3] = pack(iz + 1, ix);
    const int sti[4] = {pd.pos[0], pd.pos[1], pd.pos[2], pd.pos[3]};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (sti[i] == 0 || sti[i] == -2) continue;
      const int niz = pd.n[i] & 0xffff, nix = pd.n[i] >> 16;
      G.nd[nid(G, niz, nix)].t = tq[i];
      if (sti[i] == -1) {
        if (!heap_add(H, G, pd.n[i], tq[i], pd)) return false;
      } else {
        heap_up(H, G, pd.n[i], tq[i], pd.pos[i], pd);
      }
    }
  }
  return true;
}

__device__ __forceinline__ float bilin(float v11, float v12, float v21, float v22, float dsx,
                                       float dsz, float dnx, float dnz) {
  // bilinear (fmm.f90:1662-1682): nv(i, j), i along x, j along z
  const float nv[2][2] = {{v11, v12}, {v21, v22}};
  float b = 0.0f;
#pragma unroll
  for (int i = 1; i <= 2; ++i)
#pragma unroll
    for (int j = 1; j <= 2; ++j) {
      const float produ = (1.0f - fabsf(((float)(i - 1) * dnx - dsx) / dnx)) *
                          (1.0f - fabsf(((float)(j - 1) * dnz - dsz) / dnz));
      b = b + nv[i - 1][j - 1] * produ;
    }
  return b;
}

struct MarchArgs {
  FGrid g;
  const FPair *pairs;
  int p0, npair;
  const double *map;  // vc, (nx*ny, kmax) column-major
  long long kstride;
  const float *veln;      // coarse planes
  const float *risti_c;   // coarse column table
  const float *risti_r;   // refined column tables, kRmax per pair
  Node *cnodes;           // per slot nnx*nnz
  Node *rnodes;           // per slot kRmax*kRmax
  float *rvel;            // per slot kRmax*kRmax
  HEnt *hov;              // GH = true: hov_stride entries per listed pair
  long long hov_stride;
  const float *rcx, *rcz;
  float *ttime;
  int *rstat;
  int *pstat;
  const int *list;        // GH = true: the chunk slots to march
  int lds_cap;            // GH = false: heap entries used in LDS (<= kHeapLds)
};

// fmst_run (fmst.f90:108-332) for one station pair per workgroup.
template <bool GH>
using HeapOf = typename std::conditional<GH, GlobalHeap, LdsHeap>::type;
template <bool GH>
__device__ __forceinline__ HeapOf<GH> make_heap(HEnt *lheap, const MarchArgs &A, int sl, int maxbt) {
  if constexpr (GH) {
    return GlobalHeap{A.hov + (long long)blockIdx.x * A.hov_stride, 0, maxbt};
  } else {
    return LdsHeap{(LdsHEnt *)lheap, 0, min(maxbt, A.lds_cap)};
  }
}

// GH = false: the heap in LDS, a pair whose band outgrows it ends with
// status 9 (march it again); GH = true: the heap in global memory, over the
// pairs listed in A.list.
template <bool GH>
__global__ void __launch_bounds__(64) fmst_march(MarchArgs A) {
  __shared__ HEnt lheap[GH ? 1 : kHeapLds];
  __shared__ int ok;
  const int sl = GH ? A.list[blockIdx.x] : (int)blockIdx.x;
  if (sl >= A.npair) return;
  const FPair P = A.pairs[A.p0 + sl];
  const FGrid &g = A.g;
  const int lane = threadIdx.x;
  if (P.status != 0) {
    if (lane == 0) A.pstat[A.p0 + sl] = P.status;
    for (int r = lane; r < P.nr; r += 64) {
      A.rstat[P.r0 + r] = P.status;
      A.ttime[P.r0 + r] = 0.0f;
    }
    return;
  }
  const long long ncn = (long long)g.nnx * g.nnz;
  Node *cn = A.cnodes + (long long)sl * ncn;
  Node *rn = A.rnodes + (long long)sl * kRmax * kRmax;
  float *rv = A.rvel + (long long)sl * kRmax * kRmax;
  const double *pv = A.map + (long long)P.idx * A.kstride;
  // bsplrefine (fmm.f90:877-944) over the box, and every refined node far
  {
    const int nrxr = g.gdx * g.sgdl, nrzr = g.gdz * g.sgdl;
    const int origx = (P.vnl - 1) * g.sgdl + 1, origz = (P.vnt - 1) * g.sgdl + 1;
    const int nr = P.nrnx * P.nrnz;
    for (int q = lane; q < nr; q += 64) {
      const int idm1 = q % P.nrnz + 1, idm2 = q / P.nrnz + 1;
      const int st1 = idm1 + origz - 1, st2 = idm2 + origx - 1;
      const int i = min((st1 - 1) / nrzr + 1, g.nvz - 1), k = st1 - nrzr * (i - 1);
      const int j = min((st2 - 1) / nrxr + 1, g.nvx - 1), l = st2 - nrxr * (j - 1);
      float ub[4], vb[4];
      basis4((float)(l - 1) / (float)nrxr, ub);
      basis4((float)(k - 1) / (float)nrzr, vb);
      float sum[4];
#pragma unroll
      for (int i1 = 1; i1 <= 4; ++i1) {
        float s = 0.0f;
#pragma unroll
        for (int j1 = 1; j1 <= 4; ++j1)
          s = s + ub[j1 - 1] * (float)pv[(i - 2 + i1) * (g.nvx + 2) + (j - 2 + j1)];
        sum[i1 - 1] = vb[i1 - 1] * s;
      }
      rv[q] = sum[0] + sum[1] + sum[2] + sum[3];
      rn[q] = Node{0.0f, -1};
    }
  }
  if (lane == 0) ok = 1;
  __syncthreads();
  // the marches run on the whole wavefront with uniform control (see march)
  {
    MGrid G{rn, rv, A.risti_r + (long long)(A.p0 + sl) * kRmax, P.nrnx, P.nrnz,
            P.gorx, P.gorz, P.drnx, P.drnz, g.earth};
    auto H = make_heap<GH>(lheap, A, sl, P.maxbt);
    // the source: bilinear velocity and straight-line times to the four
    // corners of its refined cell (fmm.f90:182-207)
    const int isx = P.isxr, isz = P.iszr;
    const float scx = P.srcx, scz = P.srcz;
    float vss[2][2];
    for (int i = 1; i <= 2; ++i)
      for (int j = 1; j <= 2; ++j) vss[i - 1][j - 1] = uni(rv[nid(G, isz - 1 + j, isx - 1 + i)]);
    const float dsx = (scx - G.gox) - (float)(isx - 1) * G.dnx;
    const float dsz = (scz - G.goz) - (float)(isz - 1) * G.dnz;
    const float vsrc = bilin(vss[0][0], vss[0][1], vss[1][0], vss[1][1], dsx, dsz, G.dnx, G.dnz);
    Pending pd{{-1, -1, -1, -1}, {0, 0, 0, 0}};
    for (int i = 1; i <= 2 && ok; ++i)
      for (int j = 1; j <= 2 && ok; ++j) {
        const float ex = dsx - (float)(i - 1) * G.dnx, ez = dsz - (float)(j - 1) * G.dnz;
        const float ds = sqrtf(ex * ex + ez * ez);
        const float t = 2.0f * ds / (vss[i - 1][j - 1] + vsrc);
        rn[nid(G, isz - 1 + j, isx - 1 + i)].t = t;
        if (!heap_add(H, G, pack(isz - 1 + j, isx - 1 + i), t, pd)) ok = 0;
      }
    if (ok && !march(H, G, 1, P.vnl, P.vnr, P.vnt, P.vnb)) ok = 0;
  }
  __syncthreads();
  if (!ok) {
    // the reference's heap overflows (2), or only the LDS one (9: again
    // with the heap in global memory)
    const int code = (GH || P.maxbt <= A.lds_cap) ? 2 : 9;
    if (lane == 0) A.pstat[A.p0 + sl] = code;
    for (int r = lane; r < P.nr; r += 64) {
      A.rstat[P.r0 + r] = code;
      A.ttime[P.r0 + r] = 0.0f;
    }
    return;
  }
  // coarse grid: every node far, then the refined nodes on coarse nodes
  // (fmst.f90:240-252)
  for (long long q = lane; q < ncn; q += 64) cn[q].s = -1;
  __syncthreads();
  const int bx = P.vnr - P.vnl + 1, bz = P.vnb - P.vnt + 1;
  for (int q = lane; q < bx * bz; q += 64) {
    const int idm1 = P.vnt + q % bz, idm2 = P.vnl + q / bz;
    const int k = (idm1 - P.vnt) * g.sgdl + 1, l = (idm2 - P.vnl) * g.sgdl + 1;
    const Node r = rn[(long long)(k - 1) + (long long)(l - 1) * P.nrnz];
    Node &c = cn[(long long)(idm1 - 1) + (long long)(idm2 - 1) * g.nnz];
    if (r.s >= 0)
      c = r;
    else
      c.s = r.s;
  }
  __syncthreads();
  // alive box nodes next to far ones become close (fmst.f90:286-306)
  for (int q = lane; q < bx * bz; q += 64) {
    const int l = P.vnt + q % bz, k = P.vnl + q / bz;
    Node *c = cn + (long long)(l - 1) + (long long)(k - 1) * g.nnz;
    if (c->s == 0) {
      bool f = false;
      if (l - 1 >= 1 && c[-1].s == -1) f = true;
      if (l + 1 <= g.nnz && c[1].s == -1) f = true;
      if (k - 1 >= 1 && c[-g.nnz].s == -1) f = true;
      if (k + 1 <= g.nnx && c[g.nnz].s == -1) f = true;
      if (f) c->s = 1;
    }
  }
  __syncthreads();
  {
    MGrid G{cn, A.veln + (long long)P.plane * ncn, A.risti_c, g.nnx, g.nnz,
            g.gox, g.goz, g.dnx, g.dnz, g.earth};
    auto H = make_heap<GH>(lheap, A, sl, P.maxbt);
    Pending pd{{-1, -1, -1, -1}, {0, 0, 0, 0}};
    // the band from the close nodes, x outer, z inner (fmm.f90:188-197);
    // only box nodes can be close
    for (int i = P.vnl; i <= P.vnr && ok; ++i)
      for (int j = P.vnt; j <= P.vnb && ok; ++j) {
        const Node c = cn[nid(G, j, i)];
        if (uni(c.s) > 0 && !heap_add(H, G, pack(j, i), uni(c.t), pd)) ok = 0;
      }
    if (ok && !march(H, G, 2, 0, 0, 0, 0)) ok = 0;
  }
  __syncthreads();
  if (!ok) {
    // the reference's heap overflows (2), or only the LDS one (9: again
    // with the heap in global memory)
    const int code = (GH || P.maxbt <= A.lds_cap) ? 2 : 9;
    if (lane == 0) A.pstat[A.p0 + sl] = code;
    for (int r = lane; r < P.nr; r += 64) {
      A.rstat[P.r0 + r] = code;
      A.ttime[P.r0 + r] = 0.0f;
    }
    return;
  }
  // srtimes (fmm.f90:952-1080), a lane per receiver
  const float *vc = A.veln + (long long)P.plane * ncn;
  for (int r = lane; r < P.nr; r += 64) {
    const float rcx1 = A.rcx[P.r0 + r], rcz1 = A.rcz[P.r0 + r];
    const float scx = P.srcx, scz = P.srcz;
    int irx = (int)((rcx1 - g.gox) / g.dnx) + 1, irz = (int)((rcz1 - g.goz) / g.dnz) + 1;
    if (irx < 1 || irx > g.nnx || irz < 1 || irz > g.nnz) {
      A.rstat[P.r0 + r] = 1;
      A.ttime[P.r0 + r] = 0.0f;
      continue;
    }
    if (irx == g.nnx) irx -= 1;
    if (irz == g.nnz) irz -= 1;
    const int isx = (int)((scx - g.gox) / g.dnx) + 1, isz = (int)((scz - g.goz) / g.dnz) + 1;
    const float e1 = (scx - rcx1) * g.earth;
    const float e2 = (scz - rcz1) * g.earth * (float)sin((double)rcx1);
    const float sred = sqrtf(e1 * e1 + e2 * e2);
    int sw = 0;
    if (sred < g.dpl_sr) sw = 1;
    if (isx == irx && isz == irz) sw = 1;
    float trr;
    auto at = [&](const float *a, int iz, int ix) {
      return a[(long long)(iz - 1) + (long long)(ix - 1) * g.nnz];
    };
    if (sw == 1) {
      float drx = (scx - g.gox) - (float)(isx - 1) * g.dnx;
      float drz = (scz - g.goz) - (float)(isz - 1) * g.dnz;
      // a source on the last node line: the reference reads one node past
      // the grid there, with bilinear weight 0; the read stays inside here
      const int sx = isx, sz = isz;
      const int sx1 = min(isx + 1, g.nnx), sz1 = min(isz + 1, g.nnz);
      const float vels = bilin(at(vc, sz, sx), at(vc, sz1, sx), at(vc, sz, sx1), at(vc, sz1, sx1),
                               drx, drz, g.dnx, g.dnz);
      drx = (rcx1 - g.gox) - (float)(irx - 1) * g.dnx;
      drz = (rcz1 - g.goz) - (float)(irz - 1) * g.dnz;
      const float velr = bilin(at(vc, irz, irx), at(vc, irz + 1, irx), at(vc, irz, irx + 1),
                               at(vc, irz + 1, irx + 1), drx, drz, g.dnx, g.dnz);
      trr = 2.0f * sred / (vels + velr);
    } else {
      const float drx = (rcx1 - g.gox) - (float)(irx - 1) * g.dnx;
      const float drz = (rcz1 - g.goz) - (float)(irz - 1) * g.dnz;
      trr = 0.0f;
      for (int k = 1; k <= 2; ++k)
        for (int l = 1; l <= 2; ++l) {
          const float produ = (1.0f - fabsf(((float)(l - 1) * g.dnz - drz) / g.dnz)) *
                              (1.0f - fabsf(((float)(k - 1) * g.dnx - drx) / g.dnx));
          trr = trr + cn[(long long)(irz - 2 + l) + (long long)(irx - 2 + k) * g.nnz].t * produ;
        }
    }
    A.ttime[P.r0 + r] = trr;
    A.rstat[P.r0 + r] = 0;
  }
  if (lane == 0) A.pstat[A.p0 + sl] = 0;
}

// ---------------------------------------------------------------- rays
__device__ float delsph_dev(float c1, float l1, float c2, float l2) {
  // spherical.cpp:9-26: float variables, the trigonometry in double
  const float R = 6371.0f;
  const float dlat = c2 - c1, dlon = l2 - l1;
  const float la1 = (float)(M_PI * 0.5 - (double)c1), la2 = (float)(M_PI * 0.5 - (double)c2);
  const double s1 = sin((double)dlat * 0.5), s2 = sin((double)dlon * 0.5);
  const float a = (float)(s1 * s1 + s2 * s2 * cos((double)la1) * cos((double)la2));
  const float c = (float)(2.0 * atan2(sqrt((double)a), sqrt((double)(1 - a))));
  return R * c;
}

struct RayArgs {
  FGrid g;
  const FPair *pairs;
  int p0, npair;
  const int *rpair;       // chunk-local pair slot of each receiver
  int q0, nrcv;           // receivers q0 .. q0+nrcv-1
  int all;                // trace every receiver (else only ray-time pairs')
  const Node *cnodes, *rnodes;
  const float *velnew;    // group planes
  const float *rcx, *rcz;
  float *fdm;             // per chunk receiver, (nvx+2)*(nvz+2), x fastest
  float *tsyn;
  float *ttime;           // replaced by the ray time for ray-time pairs
  int *rstat;
};

// fmst_raypath (fmst.f90:350-976), a lane per receiver.
__global__ void __launch_bounds__(64) fmst_rays(RayArgs A) {
  const int qq = blockIdx.x * blockDim.x + threadIdx.x;
  if (qq >= A.nrcv) return;
  const int q = A.q0 + qq;
  const int sl = A.rpair[qq];
  const FPair P = A.pairs[A.p0 + sl];
  if (P.status != 0 || A.rstat[q] == 2) return;
  if (!A.all && !P.rayt) {
    A.tsyn[q] = 0.0f;
    return;
  }
  const FGrid &g = A.g;
  // the receiver's reported time (SurfaceWave.cpp:531-534, 616-618)
  auto finish = [&](float t, int status) {
    A.tsyn[q] = t;
    A.rstat[q] = status;
    if (P.rayt) A.ttime[q] = t;
  };
  const long long ncn = (long long)g.nnx * g.nnz;
  const Node *cn = A.cnodes + (long long)sl * ncn;
  const Node *rn = A.rnodes + (long long)sl * kRmax * kRmax;
  const float *vn = A.velnew + (long long)P.plane * ncn;
  float *fdm = A.fdm + (long long)qq * (g.nvx + 2) * (g.nvz + 2);
  const int fl = g.nvx + 2;
  const float scx = P.srcx, scz = P.srcz;
  const float surfrcx = A.rcx[q], surfrcz = A.rcz[q];
  const int nnxr = P.nrnx, nnzr = P.nrnz;
  const float goxr = P.gorx, gozr = P.gorz, dnxr = P.drnx, dnzr = P.drnz;
  const float gox = g.gox, goz = g.goz, dnx = g.dnx, dnz = g.dnz, earth = g.earth;
  const int isx = P.isxq, isz = P.iszq;
  const float dpl = g.dpl;
  auto tc = [&](int iz, int ix) { return cn[(long long)(iz - 1) + (long long)(ix - 1) * g.nnz].t; };
  auto nr_ = [&](int iz, int ix) { return rn[(long long)(iz - 1) + (long long)(ix - 1) * nnzr]; };
  auto vnat = [&](int iz, int ix) { return vn[(long long)(iz - 1) + (long long)(ix - 1) * g.nnz]; };
  // bilinear group velocity at a point (fmst.f90:922-955)
  auto vel_at = [&](float x, float z) {
    const int ipx = (int)((x - gox) / dnx) + 1, ipz = (int)((z - goz) / dnz) + 1;
    const float drx = (x - gox) - (float)(ipx - 1) * dnx, drz = (z - goz) - (float)(ipz - 1) * dnz;
    float vel = 0.0f;
    for (int l = 1; l <= 2; ++l)
      for (int m = 1; m <= 2; ++m) {
        float produ = 1.0f - fabsf(((float)(m - 1) * dnz - drz) / dnz);
        produ = produ * (1.0f - fabsf(((float)(l - 1) * dnx - drx) / dnx));
        if (ipz - 1 + m <= g.nnz && ipx - 1 + l <= g.nnx) vel = vel + vnat(ipz - 1 + m, ipx - 1 + l) * produ;
      }
    return vel;
  };
  auto in_refined = [&](int ipxr, int ipzr) {
    if (ipxr < 1 || ipxr >= nnxr) return 0;
    if (ipzr < 1 || ipzr >= nnzr) return 0;
    if (nr_(ipzr, ipxr).s != 0 || nr_(ipzr + 1, ipxr).s != 0) return 0;
    if (nr_(ipzr, ipxr + 1).s != 0 || nr_(ipzr + 1, ipxr + 1).s != 0) return 0;
    return 1;
  };
  int ipx = (int)((surfrcx - gox) / dnx) + 1, ipz = (int)((surfrcz - goz) / dnz) + 1;
  if (ipx < 1 || ipx >= g.nnx || ipz < 1 || ipz >= g.nnz) {
    finish(0.0f, 1);
    return;
  }
  float x0 = surfrcx, z0 = surfrcz;  // rgx(j), rgz(j)
  float tsyn = 0.0f;
  int sw = 0;
  {
    const float e1 = (scx - x0) * earth, e2 = (scz - z0) * earth * (float)sin((double)x0);
    if (sqrtf(e1 * e1 + e2 * e2) < 2.0f * dpl) sw = 1;
  }
  int ipxr = (int)((surfrcx - goxr) / dnxr) + 1, ipzr = (int)((surfrcz - gozr) / dnzr) + 1;
  int igref = in_refined(ipxr, ipzr);
  if (sw == 0 && igref == 1 && ipxr == isx && ipzr == isz) sw = 1;
  // the group time of a segment, in the order of the reference's final loop
  auto seg_time = [&](float xa, float za, float xb, float zb) {
    const float dinc = delsph_dev(xa, za, xb, zb);
    const float velo = vel_at(xa, za), vel = vel_at(xb, zb);
    return dinc * 0.5f * (1.0f / vel + 1.0f / velo);
  };
  if (sw == 1) {
    tsyn = tsyn + seg_time(x0, z0, scx, scz);
    finish(tsyn, 0);
    return;
  }
  const int maxrp = g.nnx * g.nnz;
  for (int jstep = 1; jstep <= maxrp; ++jstep) {
    float dtx, dtz;
    const float sx0 = (float)sin((double)x0);
    if (igref == 1) {
      dtx = nr_(ipzr, ipxr + 1).t - nr_(ipzr, ipxr).t;
      dtx = dtx + nr_(ipzr + 1, ipxr + 1).t - nr_(ipzr + 1, ipxr).t;
      dtx = dtx / (2.0f * earth * dnxr);
      dtz = nr_(ipzr + 1, ipxr).t - nr_(ipzr, ipxr).t;
      dtz = dtz + nr_(ipzr + 1, ipxr + 1).t - nr_(ipzr, ipxr + 1).t;
      dtz = dtz / (2.0f * earth * sx0 * dnzr);
    } else {
      dtx = tc(ipz, ipx + 1) - tc(ipz, ipx);
      dtx = dtx + tc(ipz + 1, ipx + 1) - tc(ipz + 1, ipx);
      dtx = dtx / (2.0f * earth * dnx);
      dtz = tc(ipz + 1, ipx) - tc(ipz, ipx);
      dtz = dtz + tc(ipz + 1, ipx + 1) - tc(ipz, ipx + 1);
      dtz = dtz / (2.0f * earth * sx0 * dnz);
    }
    const float rd1 = sqrtf(dtx * dtx + dtz * dtz);
    float x1 = x0 - dpl * dtx / (earth * rd1);
    float z1 = z0 - dpl * dtz / (earth * sx0 * rd1);
    const int ipxo = ipx, ipzo = ipz;
    ipxr = (int)((x1 - goxr) / dnxr) + 1;
    ipzr = (int)((z1 - gozr) / dnzr) + 1;
    igref = in_refined(ipxr, ipzr);
    ipx = (int)((x1 - gox) / dnx) + 1;
    ipz = (int)((z1 - goz) / dnz) + 1;
    {
      const float e1 = (scx - x1) * earth, e2 = (scz - z1) * earth * (float)sin((double)x1);
      sw = 0;
      if (sqrtf(e1 * e1 + e2 * e2) < 2.0f * dpl) sw = 1;
    }
    if (sw == 0 && igref == 1 && ipxr == isx && ipzr == isz) sw = 1;
    if (ipx < 1) {
      x1 = gox;
      ipx = 1;
    }
    if (ipx >= g.nnx) {
      x1 = gox + (float)(g.nnx - 1) * dnx;
      ipx = g.nnx - 1;
    }
    if (ipz < 1) {
      z1 = goz;
      ipz = 1;
    }
    if (ipz >= g.nnz) {
      z1 = goz + (float)(g.nnz - 1) * dnz;
      ipz = g.nnz - 1;
    }
    // Fréchet weights of the segment (x0, z0) -> (x1, z1) (fmst.f90:700-848)
    const int ivx = (ipx - 1) / g.gdx + 1, ivz = (ipz - 1) / g.gdz + 1;
    const int ivxo = (ipxo - 1) / g.gdx + 1, ivzo = (ipzo - 1) / g.gdz + 1;
    int nhp = 0, chp[3];
    float vrat[3];
    if (ivx != ivxo) {
      nhp = nhp + 1;
      const float xi = (ivx > ivxo) ? gox + (float)(ivx - 1) * g.dvx : gox + (float)ivx * g.dvx;
      vrat[nhp - 1] = (xi - x0) / (x1 - x0);
      chp[nhp - 1] = 1;
    }
    if (ivz != ivzo) {
      nhp = nhp + 1;
      const float zi = (ivz > ivzo) ? goz + (float)(ivz - 1) * g.dvz : goz + (float)ivz * g.dvz;
      const float r = (zi - z0) / (z1 - z0);
      if (nhp == 1) {
        vrat[0] = r;
        chp[0] = 2;
      } else if (r >= vrat[0]) {
        vrat[1] = r;
        chp[1] = 2;
      } else {
        vrat[1] = vrat[0];
        chp[1] = chp[0];
        vrat[0] = r;
        chp[0] = 2;
      }
    }
    nhp = nhp + 1;
    vrat[nhp - 1] = 1.0f;
    chp[nhp - 1] = 0;
    float vel = 0.0f;
    {
      const float drx = (x0 - gox) - (float)(ipxo - 1) * dnx, drz = (z0 - goz) - (float)(ipzo - 1) * dnz;
      for (int l = 1; l <= 2; ++l)
        for (int m = 1; m <= 2; ++m) {
          float produ = 1.0f - fabsf(((float)(m - 1) * dnz - drz) / dnz);
          produ = produ * (1.0f - fabsf(((float)(l - 1) * dnx - drx) / dnx));
          if (ipzo - 1 + m <= g.nnz && ipxo - 1 + l <= g.nnx) vel = vel + vnat(ipzo - 1 + m, ipxo - 1 + l) * produ;
        }
    }
    float vi[4], wi[4];
    basis4(((x0 - gox) - (float)(ivxo - 1) * g.dvx) / g.dvx, vi);
    basis4(((z0 - goz) - (float)(ivzo - 1) * g.dvz) / g.dvz, wi);
    int ivxt = ivxo, ivzt = ivzo;
    for (int k = 1; k <= nhp; ++k) {
      const float velo = vel;
      float vio[4], wio[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        vio[i] = vi[i];
        wio[i] = wi[i];
      }
      if (k > 1) {
        if (chp[k - 2] == 1)
          ivxt = ivx;
        else if (chp[k - 2] == 2)
          ivzt = ivz;
      }
      const float rigz = z0 + vrat[k - 1] * (z1 - z0);
      const float rigx = x0 + vrat[k - 1] * (x1 - x0);
      const int ipxt = (int)((rigx - gox) / dnx) + 1, ipzt = (int)((rigz - goz) / dnz) + 1;
      float drx = (rigx - gox) - (float)(ipxt - 1) * dnx, drz = (rigz - goz) - (float)(ipzt - 1) * dnz;
      vel = 0.0f;
      for (int m = 1; m <= 2; ++m)
        for (int n = 1; n <= 2; ++n) {
          float produ = 1.0f - fabsf(((float)(n - 1) * dnz - drz) / dnz);
          produ = produ * (1.0f - fabsf(((float)(m - 1) * dnx - drx) / dnx));
          if (ipzt - 1 + n <= g.nnz && ipxt - 1 + m <= g.nnx) vel = vel + vnat(ipzt - 1 + n, ipxt - 1 + m) * produ;
        }
      drx = (rigx - gox) - (float)(ivxt - 1) * g.dvx;
      drz = (rigz - goz) - (float)(ivzt - 1) * g.dvz;
      basis4(drx / g.dvx, vi);
      basis4(drz / g.dvz, wi);
      const float dinc = (k == 1) ? vrat[0] * dpl : (vrat[k - 1] - vrat[k - 2]) * dpl;
      for (int l = 1; l <= 4; ++l)
        for (int m = 1; m <= 4; ++m) {
          float r1 = vi[m - 1] * wi[l - 1] / (vel * vel);
          const float r2 = vio[m - 1] * wio[l - 1] / (velo * velo);
          r1 = -((r1 + r2) * dinc / 2.0f);
          float &f = fdm[(long long)(ivzt - 2 + l) * fl + (ivxt - 2 + m)];
          f = r1 + f;
        }
    }
    tsyn = tsyn + seg_time(x0, z0, x1, z1);
    x0 = x1;
    z0 = z1;
    if (sw == 1) {
      tsyn = tsyn + seg_time(x0, z0, scx, scz);
      finish(tsyn, 0);
      return;
    }
  }
  finish(0.0f, 3);  // the ray never reached the source
}

// nonzero weights per receiver (count), then their cells and weights in
// increasing n = ix + iy*nx (fill); a wavefront per receiver
__global__ void __launch_bounds__(64)
fmst_ray_count(int nrcv, int ncell, const float *fdm, long long *cnt) {
  const int q = blockIdx.x;
  if (q >= nrcv) return;
  const float *f = fdm + (long long)q * ncell;
  int c = 0;
  for (int i = threadIdx.x; i < ncell; i += 64) c += (f[i] != 0.0f) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (threadIdx.x == 0) cnt[q] = c;
}

__global__ void __launch_bounds__(64)
fmst_ray_fill(int nrcv, int ncell, const float *fdm, const long long *off, int *cell, float *w) {
  const int q = blockIdx.x;
  if (q >= nrcv) return;
  const float *f = fdm + (long long)q * ncell;
  long long o = off[q];
  const int lane = threadIdx.x;
  for (int i0 = 0; i0 < ncell; i0 += 64) {
    const int i = i0 + lane;
    const float v = (i < ncell) ? f[i] : 0.0f;
    const bool nz = v != 0.0f;
    const unsigned long long m = __ballot(nz);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (nz) {
      cell[o + below] = i;
      w[o + below] = v;
    }
    o += __popcll(m);
  }
}

__global__ void fmst_zero(long long n, float *p) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

}  // namespace

// ---------------------------------------------------------------- host API
struct jdsg_fmst {
  FGrid g;
  int npair = 0, nrcv = 0;
  std::vector<FPair> pairs;
  std::vector<int> rpair;                 // pair of each receiver
  std::vector<int> plane_idx;             // period index of each plane
  std::vector<int> ray_time;              // per pair: receivers take the ray time
  FPair *d_pairs = nullptr;
  float *d_rcx = nullptr, *d_rcz = nullptr;
  float *d_risti_c = nullptr, *d_risti_r = nullptr;
  int *d_plane_idx = nullptr;
  long long hov_max = 0;                  // largest heap past the LDS capacity (0: none)
  // results of the last run
  std::vector<long long> ray_cnt;         // per receiver nonzero weights (rays kept)
  bool have_rays = false;
  float *d_fdm_all = nullptr;             // dense weights of every receiver (rays kept)
  // device work buffers kept between runs (an inversion calls run every
  // iteration): slot -> (pointer, bytes)
  std::vector<std::pair<void *, size_t>> bufs;
  void *buf(int slot, size_t bytes) {
    if ((int)bufs.size() <= slot) bufs.resize(slot + 1, {nullptr, 0});
    auto &b = bufs[slot];
    if (b.second < bytes) {
      (void)hipFree(b.first);
      b = {nullptr, 0};
      if (hipMalloc(&b.first, std::max<size_t>(bytes, 1)) != hipSuccess) return nullptr;
      b.second = bytes;
    }
    return b.first;
  }
  ~jdsg_fmst() {
    for (auto &b : bufs) (void)hipFree(b.first);
    (void)hipFree(d_pairs);
    (void)hipFree(d_rcx);
    (void)hipFree(d_rcz);
    (void)hipFree(d_risti_c);
    (void)hipFree(d_risti_r);
    (void)hipFree(d_plane_idx);
  }
};

namespace {

// fmst_init (fmst.f90:9-82) in float, as the reference computes it
FGrid make_grid(int nx, int ny, float goxdf, float gozdf, float dvxdf, float dvzdf) {
  FGrid g{};
  g.nx = nx;
  g.ny = ny;
  g.gdx = 8;
  g.gdz = 8;
  g.sgdl = 8;
  g.earth = 6371.0f;
  g.snb = 0.5f;
  g.nvx = nx - 2;
  g.nvz = ny - 2;
  const float goxd = goxdf - dvxdf, gozd = gozdf + dvzdf;
  g.dvx = dvxdf * kPi / 180.0f;
  g.dvz = dvzdf * kPi / 180.0f;
  g.gox = (90.0f - goxd) * kPi / 180.0f;
  g.goz = gozd * kPi / 180.0f;
  g.nnx = (g.nvx - 1) * g.gdx + 1;
  g.nnz = (g.nvz - 1) * g.gdz + 1;
  g.dnx = g.dvx / (float)g.gdx;
  g.dnz = g.dvz / (float)g.gdz;
  // srtimes / fmst_raypath step (fmm.f90:1027-1031, fmst.f90:459-464)
  float dpl = g.dnx * g.earth;
  float rd1 = g.dnz * g.earth * std::sin(g.gox);
  if (rd1 < dpl) dpl = rd1;
  rd1 = g.dnz * g.earth * std::sin(g.gox + (float)(g.nnx - 1) * g.dnx);
  if (rd1 < dpl) dpl = rd1;
  g.dpl_sr = dpl;
  g.dpl = 0.5f * dpl;
  return g;
}

// fmst_run's refined-grid constants (fmst.f90:140-197) and heap size
void make_pair(const FGrid &g, float srcx, float srcz, FPair &P, std::vector<float> &risti) {
  const int sgs = 8;
  P.status = 0;
  const float x = srcx, z = srcz;
  int isx = (int)((x - g.gox) / g.dnx) + 1, isz = (int)((z - g.goz) / g.dnz) + 1;
  if (isx < 1 || isx > g.nnx || isz < 1 || isz > g.nnz) {
    P.status = 1;
    return;
  }
  if (isx == g.nnx) isx -= 1;
  if (isz == g.nnz) isz -= 1;
  P.vnl = std::max(isx - sgs, 1);
  P.vnr = std::min(isx + sgs, g.nnx);
  P.vnt = std::max(isz - sgs, 1);
  P.vnb = std::min(isz + sgs, g.nnz);
  P.nrnx = (P.vnr - P.vnl) * g.sgdl + 1;
  P.nrnz = (P.vnb - P.vnt) * g.sgdl + 1;
  P.drnx = g.dvx / (float)(g.gdx * g.sgdl);
  P.drnz = g.dvz / (float)(g.gdz * g.sgdl);
  P.gorx = g.gox + g.dnx * (float)(P.vnl - 1);
  P.gorz = g.goz + g.dnz * (float)(P.vnt - 1);
  const int idm1 = std::max(P.nrnx, g.nnx), idm2 = std::max(P.nrnz, g.nnz);
  P.maxbt = (P.nrnx > g.nnx || P.nrnz > g.nnz) ? (int)std::round(g.snb * (float)idm1 * (float)idm2)
                                              : (int)std::round(g.snb * (float)g.nnx * (float)g.nnz);
  // travel (urg = 1) on the refined grid (fmm.f90:162-176)
  int a = (int)((x - P.gorx) / P.drnx) + 1, b = (int)((z - P.gorz) / P.drnz) + 1;
  P.isxq = a;
  P.iszq = b;
  if (a < 1 || a > P.nrnx || b < 1 || b > P.nrnz) {
    P.status = 1;
    return;
  }
  if (a == P.nrnx) a -= 1;
  if (b == P.nrnz) b -= 1;
  P.isxr = a;
  P.iszr = b;
  for (int ix = 1; ix <= P.nrnx; ++ix)
    risti.push_back(g.earth * std::sin(P.gorx + (float)(ix - 1) * P.drnx));
  risti.resize(risti.size() + (kRmax - P.nrnx), 0.0f);
}

template <class T>
int upload(T **d, const T *h, size_t n) {
  FCHK(hipMalloc((void **)d, sizeof(T) * std::max<size_t>(n, 1)));
  if (n) FCHK(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" const char *jdsg_fmst_last_error(void) { return g_err.c_str(); }

extern "C" int jdsg_fmst_create(int nx, int ny, float lat0, float lon0, float dlat, float dlon,
                                int npair, const int *idx, const float *src_colat,
                                const float *src_lon, const int *rcv_off, const float *rcv_colat,
                                const float *rcv_lon, const int *ray_time, jdsg_fmst **out) {
  if (!out) return fail("fmst_create: null handle pointer");
  *out = nullptr;
  if (nx < 4 || ny < 4) return fail("fmst_create: need nx, ny >= 4");
  if (npair < 0 || (npair > 0 && (!idx || !src_colat || !src_lon || !rcv_off)))
    return fail("fmst_create: missing pair arrays");
  auto *h = new jdsg_fmst();
  std::unique_ptr<jdsg_fmst> own(h);
  h->g = make_grid(nx, ny, lat0, lon0, dlat, dlon);
  if ((long long)h->g.nnx >= 65536 || (long long)h->g.nnz >= 65536)
    return fail("fmst_create: propagation grid above 65535 nodes per axis");
  h->npair = npair;
  h->nrcv = npair ? rcv_off[npair] : 0;
  if (npair && rcv_off[0] != 0) return fail("fmst_create: rcv_off[0] must be 0");
  if (h->nrcv && (!rcv_colat || !rcv_lon)) return fail("fmst_create: missing receiver arrays");
  std::vector<float> risti_r;
  std::vector<int> plane_of(0);
  h->pairs.resize(npair);
  h->ray_time.assign(npair, 0);
  for (int p = 0; p < npair; ++p) {
    if (rcv_off[p + 1] < rcv_off[p]) return fail("fmst_create: rcv_off decreases");
    if (idx[p] < 0) return fail("fmst_create: negative period index");
    FPair &P = h->pairs[p];
    std::memset(&P, 0, sizeof(P));
    P.idx = idx[p];
    P.r0 = rcv_off[p];
    P.nr = rcv_off[p + 1] - rcv_off[p];
    P.srcx = src_colat[p];
    P.srcz = src_lon[p];
    const size_t before = risti_r.size();
    make_pair(h->g, P.srcx, P.srcz, P, risti_r);
    risti_r.resize(before + kRmax, 0.0f);
    if (P.maxbt > kHeapLds) h->hov_max = std::max<long long>(h->hov_max, P.maxbt);
    auto it = std::find(h->plane_idx.begin(), h->plane_idx.end(), P.idx);
    if (it == h->plane_idx.end()) {
      P.plane = (int)h->plane_idx.size();
      h->plane_idx.push_back(P.idx);
    } else {
      P.plane = (int)(it - h->plane_idx.begin());
    }
    h->ray_time[p] = ray_time ? (ray_time[p] != 0) : 0;
    P.rayt = h->ray_time[p];
    for (int r = P.r0; r < P.r0 + P.nr; ++r) h->rpair.push_back(p);
  }
  std::vector<float> risti_c(h->g.nnx);
  for (int ix = 1; ix <= h->g.nnx; ++ix)
    risti_c[ix - 1] = h->g.earth * std::sin(h->g.gox + (float)(ix - 1) * h->g.dnx);
  if (upload(&h->d_pairs, h->pairs.data(), h->pairs.size()) ||
      upload(&h->d_rcx, rcv_colat, (size_t)h->nrcv) || upload(&h->d_rcz, rcv_lon, (size_t)h->nrcv) ||
      upload(&h->d_risti_c, risti_c.data(), risti_c.size()) ||
      upload(&h->d_risti_r, risti_r.data(), risti_r.size()) ||
      upload(&h->d_plane_idx, h->plane_idx.data(), h->plane_idx.size()))
    return 1;
  *out = own.release();
  return 0;
}

extern "C" int jdsg_fmst_sizes(const jdsg_fmst *h, int *npair, int *nrcv, int *nplane) {
  if (!h) return fail("fmst_sizes: null handle");
  if (npair) *npair = h->npair;
  if (nrcv) *nrcv = h->nrcv;
  if (nplane) *nplane = (int)h->plane_idx.size();
  return 0;
}

// One pass over all pairs.  mode 0: SurfTime::travel_time (rays only for the
// pairs whose receivers take the ray time); mode 1: frechet_matrix /
// compute_grad (rays for every receiver, weights kept for jdsg_fmst_rays).
extern "C" int jdsg_fmst_run(jdsg_fmst *h, const double *vc, const double *vout, int kmax,
                             int mode, float *ttime, float *tsyn, int *rstatus, int *pstatus,
                             void *stream) {
  if (!h) return fail("fmst_run: null handle");
  if (mode != 0 && mode != 1) return fail("fmst_run: mode must be 0 or 1");
  const FGrid &g = h->g;
  for (int p : h->plane_idx)
    if (p >= kmax) return fail("fmst_run: a pair's period index is not below kmax");
  hipStream_t st = (hipStream_t)stream;
  h->have_rays = false;
  h->d_fdm_all = nullptr;
  if (h->npair == 0) return 0;
  if (!vc || !vout || !ttime || !tsyn || !rstatus || !pstatus)
    return fail("fmst_run: null output or map pointer");
  const long long nxy = (long long)g.nx * g.ny, ncn = (long long)g.nnx * g.nnz;
  const int nplane = (int)h->plane_idx.size();
  const long long ncell = (long long)(g.nvx + 2) * (g.nvz + 2);
  // device memory: planes, then as many pair slots as 0.5 of the free memory takes
  float *d_veln = (float *)h->buf(0, sizeof(float) * ncn * nplane);
  float *d_velnew = (float *)h->buf(1, sizeof(float) * ncn * nplane);
  if (!d_veln || !d_velnew) return fail("fmst_run: device allocation");
  struct Free {
    std::vector<void *> p;
    ~Free() {
      for (void *q : p) (void)hipFree(q);
    }
  } fr;
  {
    dim3 grid((unsigned)((ncn + 255) / 256), (unsigned)nplane);
    hipLaunchKernelGGL(fmst_gridder, grid, dim3(256), 0, st, g, nplane, h->d_plane_idx, vc, nxy,
                       d_veln);
    hipLaunchKernelGGL(fmst_gridder, grid, dim3(256), 0, st, g, nplane, h->d_plane_idx, vout, nxy,
                       d_velnew);
    FCHK(hipGetLastError());
  }
  int maxnr = 0;
  for (const auto &P : h->pairs) maxnr = std::max(maxnr, P.nr);
  const double slot_bytes = (double)ncn * sizeof(Node) + (double)kRmax * kRmax * (sizeof(Node) + 4);
  const double rcv_bytes = (double)ncell * 4 + 16;
  size_t fre = 0, tot = 0;
  double budget = (hipMemGetInfo(&fre, &tot) == hipSuccess) ? 0.5 * (double)fre : 4e9;
  if (const char *b = getenv("JDSG_FMST_BUDGET_GB")) budget = atof(b) * 1e9;
  const bool keep = (mode == 1);
  if (keep) {
    // every traced receiver's dense weights stay for jdsg_fmst_rays
    const double need = (double)h->nrcv * ncell * 4;
    if (need > 0.5 * budget) return fail("fmst_run: dense ray weights exceed the memory budget");
    h->d_fdm_all = (float *)h->buf(2, sizeof(float) * (size_t)h->nrcv * ncell);
    if (!h->d_fdm_all) return fail("fmst_run: device allocation");
    budget -= need;
  }
  const double per_pair = slot_bytes + (keep ? 0.0 : maxnr * rcv_bytes);
  int chunk = (int)std::min<double>(h->npair, std::max(1.0, budget / per_pair));
  Node *d_cn = nullptr, *d_rn = nullptr;
  float *d_rv = nullptr, *d_fdm = nullptr;
  int *d_rpair = nullptr;
  d_cn = (Node *)h->buf(3, sizeof(Node) * ncn * chunk);
  d_rn = (Node *)h->buf(4, sizeof(Node) * kRmax * kRmax * chunk);
  d_rv = (float *)h->buf(5, sizeof(float) * kRmax * kRmax * chunk);
  if (!d_cn || !d_rn || !d_rv) return fail("fmst_run: device allocation");

  if (!keep) {
    d_fdm = (float *)h->buf(6, sizeof(float) * ncell * std::max(1, maxnr) * chunk);
    if (!d_fdm) return fail("fmst_run: device allocation");
  }
  d_rpair = (int *)h->buf(7, sizeof(int) * std::max(1, h->nrcv));
  if (!d_rpair) return fail("fmst_run: device allocation");
  std::vector<int> rp_local(h->nrcv);
  for (int p0 = 0; p0 < h->npair; p0 += chunk) {
    const int np = std::min(chunk, h->npair - p0);
    MarchArgs M{g, h->d_pairs, p0, np, vc, nxy, d_veln, h->d_risti_c, h->d_risti_r, d_cn, d_rn,
                d_rv, nullptr, 0, h->d_rcx, h->d_rcz, ttime, rstatus, pstatus, nullptr, kHeapLds};
    // a smaller LDS heap sends more pairs to the global-memory heap (tests)
    if (const char *e = getenv("JDSG_FMST_LDS_HEAP")) M.lds_cap = std::max(4, std::min(kHeapLds, atoi(e)));
    hipLaunchKernelGGL(fmst_march<false>, dim3(np), dim3(64), 0, st, M);
    FCHK(hipGetLastError());
    // pairs whose band outgrew the LDS heap: again with the heap in global memory
    if (h->hov_max > 0 || M.lds_cap < kHeapLds) {
      std::vector<int> ps(np);
      FCHK(hipMemcpyAsync(ps.data(), pstatus + p0, sizeof(int) * np, hipMemcpyDeviceToHost, st));
      FCHK(hipStreamSynchronize(st));
      std::vector<int> again;
      for (int i = 0; i < np; ++i)
        if (ps[i] == 9) again.push_back(i);
      if (!again.empty()) {
        long long stride = 1;
        for (int i : again) stride = std::max<long long>(stride, h->pairs[p0 + i].maxbt);
        HEnt *d_hov = nullptr;
        int *d_list = nullptr;
        FCHK(hipMalloc((void **)&d_hov, sizeof(HEnt) * stride * again.size()));
        fr.p.push_back(d_hov);
        FCHK(hipMalloc((void **)&d_list, sizeof(int) * again.size()));
        fr.p.push_back(d_list);
        FCHK(hipMemcpyAsync(d_list, again.data(), sizeof(int) * again.size(), hipMemcpyHostToDevice,
                            st));
        MarchArgs MG = M;
        MG.hov = d_hov;
        MG.hov_stride = stride;
        MG.list = d_list;
        hipLaunchKernelGGL(fmst_march<true>, dim3((unsigned)again.size()), dim3(64), 0, st, MG);
        FCHK(hipGetLastError());
      }
    }
    // the chunk's traced receivers, contiguous from the first pair's
    const int q0 = h->pairs[p0].r0;
    const int q1 = h->pairs[p0 + np - 1].r0 + h->pairs[p0 + np - 1].nr;
    if (q1 > q0) {
      for (int q = q0; q < q1; ++q) rp_local[q - q0] = h->rpair[q] - p0;
      FCHK(hipMemcpyAsync(d_rpair, rp_local.data(), sizeof(int) * (q1 - q0), hipMemcpyHostToDevice,
                          st));
      float *fd = keep ? h->d_fdm_all + (long long)q0 * ncell : d_fdm;
      const long long nz = (long long)(q1 - q0) * ncell;
      hipLaunchKernelGGL(fmst_zero, dim3((unsigned)std::min<long long>((nz + 255) / 256, 65536)),
                         dim3(256), 0, st, nz, fd);
      RayArgs R{g, h->d_pairs, p0, np, d_rpair, q0, q1 - q0, mode, d_cn, d_rn, d_velnew,
                h->d_rcx, h->d_rcz, fd, tsyn, ttime, rstatus};
      hipLaunchKernelGGL(fmst_rays, dim3((q1 - q0 + 63) / 64), dim3(64), 0, st, R);
      FCHK(hipGetLastError());
      FCHK(hipStreamSynchronize(st));  // rp_local is reused by the next chunk
    }
  }
  FCHK(hipStreamSynchronize(st));
  h->have_rays = keep;
  return 0;
}

// The sparse ray weights of the last mode-1 run: receiver q's nonzero cells
// n = ix + iy*nx in increasing order at ray_off[q] .. ray_off[q+1] of
// ray_cell / ray_w (device arrays; ray_off has nrcv + 1 entries).  Fails,
// with *nnz set to the size needed, when that exceeds cap.
extern "C" int jdsg_fmst_rays(jdsg_fmst *h, long long *ray_off, int *ray_cell, float *ray_w,
                              long long cap, long long *nnz, void *stream) {
  if (!h) return fail("fmst_rays: null handle");
  if (!h->have_rays) return fail("fmst_rays: no mode-1 run to take the rays from");
  hipStream_t st = (hipStream_t)stream;
  const FGrid &g = h->g;
  const int ncell = (g.nvx + 2) * (g.nvz + 2);
  const int n = h->nrcv;
  long long *d_cnt = nullptr;
  FCHK(hipMalloc((void **)&d_cnt, sizeof(long long) * std::max(n, 1)));
  std::unique_ptr<long long, decltype(&hipFree)> own(d_cnt, &hipFree);
  if (n) hipLaunchKernelGGL(fmst_ray_count, dim3(n), dim3(64), 0, st, n, ncell, h->d_fdm_all, d_cnt);
  FCHK(hipGetLastError());
  std::vector<long long> cnt(n), off(n + 1, 0);
  FCHK(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(long long) * n, hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  for (int q = 0; q < n; ++q) off[q + 1] = off[q] + cnt[q];
  if (nnz) *nnz = off[n];
  if (off[n] > cap) return fail("fmst_rays: the rays need more than cap entries");
  if (!ray_off || (off[n] > 0 && (!ray_cell || !ray_w))) return fail("fmst_rays: null output");
  FCHK(hipMemcpyAsync(ray_off, off.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, st));
  if (n) hipLaunchKernelGGL(fmst_ray_fill, dim3(n), dim3(64), 0, st, n, ncell, h->d_fdm_all, ray_off,
                            ray_cell, ray_w);
  FCHK(hipGetLastError());
  FCHK(hipStreamSynchronize(st));
  return 0;
}

extern "C" void jdsg_fmst_destroy(jdsg_fmst *h) { delete h; }

// ---------------------------------------------------------------- SurfTime
// The three SurfTime methods that run FMST per station pair, as one call
// each on device buffers: the dispersion maps (and depth kernels) of the
// model, the pairs' traveltimes and rays, then the Fréchet rows or the
// misfit gradient.
namespace {

struct StDims {
  int nx, ny, nz, kmax;
};

int st_check(jdsg_surftime *st, jdsg_fmst *fm, StDims *d) {
  if (!st || !fm) return fail("surftime/fmst: null handle");
  if (jdsg_surftime_dims(st, &d->nx, &d->ny, &d->nz, &d->kmax))
    return fail(std::string("surftime: ") + jdsg_last_error());
  if (d->nx != fm->g.nx || d->ny != fm->g.ny)
    return fail("surftime/fmst: the FMST grid is not the model's nx, ny");
  for (int p : fm->plane_idx)
    if (p >= d->kmax) return fail("surftime/fmst: a pair's period index is not below kmax");
  return 0;
}

struct DBuf {
  std::vector<void *> p;
  template <class T>
  T *get(size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T *)q;
  }
  ~DBuf() {
    for (void *q : p) (void)hipFree(q);
  }
};

// period index of every receiver's row (the pair's), on the device
int *row_index(jdsg_fmst *fm, DBuf &B) {
  std::vector<int> ri(fm->nrcv);
  for (int q = 0; q < fm->nrcv; ++q) ri[q] = fm->pairs[fm->rpair[q]].idx;
  int *d = B.get<int>(ri.size());
  if (d && !ri.empty() && hipMemcpy(d, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  return d;
}

}  // namespace

// SurfTime::travel_time (SurfaceWave.cpp:509-541): get_2d_map, then each
// pair's receiver times (ray times for Rg/Lg pairs) into data[nrcv].
extern "C" int jdsg_surftime_travel_time(jdsg_surftime *st, jdsg_fmst *fm, const float *vs,
                                         float *data, int *rstatus, void *stream) {
  StDims d;
  if (st_check(st, fm, &d)) return 1;
  DBuf B;
  const size_t nk = (size_t)d.nx * d.ny * d.kmax;
  double *vc = B.get<double>(nk), *vout = B.get<double>(nk);
  int *mst = B.get<int>(nk), *pst = B.get<int>(fm->npair);
  float *tsyn = B.get<float>(fm->nrcv);
  if (!vc || !vout || !mst || !pst || !tsyn) return fail("surftime_travel_time: device allocation");
  if (jdsg_surftime_map(st, vs, vc, vout, mst, stream))
    return fail(std::string("surftime_map: ") + jdsg_last_error());
  return jdsg_fmst_run(fm, vc, vout, d.kmax, 0, data, tsyn, rstatus, pst, stream);
}

// SurfTime::frechet_matrix (SurfaceWave.cpp:551-676): get_1d_kernel, every
// receiver's ray, data[nrcv] and the rows in CSR (device row_off[nrcv+1],
// col, val: the layout jdsg_frechet_rows writes, rows in data order).
extern "C" int jdsg_surftime_frechet(jdsg_surftime *st, jdsg_fmst *fm, const float *vs,
                                     float *data, int *rstatus, long long *row_off, int *col,
                                     float *val, long long cap, long long *nnz, void *stream) {
  StDims d;
  if (st_check(st, fm, &d)) return 1;
  DBuf B;
  const size_t nk = (size_t)d.nx * d.ny * d.kmax;
  double *vc = B.get<double>(nk), *vout = B.get<double>(nk), *ker = B.get<double>(nk * d.nz);
  int *mst = B.get<int>(nk), *pst = B.get<int>(fm->npair);
  float *tsyn = B.get<float>(fm->nrcv);
  if (!vc || !vout || !ker || !mst || !pst || !tsyn)
    return fail("surftime_frechet: device allocation");
  if (hipMemsetAsync(ker, 0, sizeof(double) * nk * d.nz, (hipStream_t)stream) != hipSuccess)
    return fail("surftime_frechet: memset");
  if (jdsg_surftime_kernel(st, vs, vc, vout, ker, mst, stream))
    return fail(std::string("surftime_kernel: ") + jdsg_last_error());
  if (jdsg_fmst_run(fm, vc, vout, d.kmax, 1, data, tsyn, rstatus, pst, stream)) return 1;
  long long nray = 0;
  jdsg_fmst_rays(fm, nullptr, nullptr, nullptr, 0, &nray, stream);
  long long *roff = B.get<long long>((size_t)fm->nrcv + 1);
  int *cell = B.get<int>((size_t)nray);
  float *w = B.get<float>((size_t)nray);
  int *ridx = row_index(fm, B);
  if (!roff || !cell || !w || !ridx) return fail("surftime_frechet: device allocation");
  if (jdsg_fmst_rays(fm, roff, cell, w, nray, &nray, stream)) return 1;
  if (jdsg_frechet_rows(d.nx, d.ny, d.nz, d.kmax, ker, fm->nrcv, ridx, roff, cell, w, row_off, col,
                        val, cap, nnz, stream))
    return fail(std::string("frechet_rows: ") + jdsg_last_error());
  return 0;
}

// SurfTime::compute_grad (SurfaceWave.cpp:695-781): get_1d_kernel, every
// receiver's ray, data[nrcv], and grad[(nx-2)(ny-2)(nz-1)] = sum over rows
// of row * (data - obst), times vs at the cell.
extern "C" int jdsg_surftime_compute_grad(jdsg_surftime *st, jdsg_fmst *fm, const float *vs,
                                          const float *obst, float *data, float *grad,
                                          int *rstatus, void *stream) {
  StDims d;
  if (st_check(st, fm, &d)) return 1;
  DBuf B;
  const size_t nk = (size_t)d.nx * d.ny * d.kmax;
  double *vc = B.get<double>(nk), *vout = B.get<double>(nk), *ker = B.get<double>(nk * d.nz);
  int *mst = B.get<int>(nk), *pst = B.get<int>(fm->npair);
  float *tsyn = B.get<float>(fm->nrcv), *resid = B.get<float>(fm->nrcv);
  if (!vc || !vout || !ker || !mst || !pst || !tsyn || !resid)
    return fail("surftime_compute_grad: device allocation");
  hipStream_t s = (hipStream_t)stream;
  if (hipMemsetAsync(ker, 0, sizeof(double) * nk * d.nz, s) != hipSuccess)
    return fail("surftime_compute_grad: memset");
  if (jdsg_surftime_kernel(st, vs, vc, vout, ker, mst, stream))
    return fail(std::string("surftime_kernel: ") + jdsg_last_error());
  if (jdsg_fmst_run(fm, vc, vout, d.kmax, 1, data, tsyn, rstatus, pst, stream)) return 1;
  long long nray = 0;
  jdsg_fmst_rays(fm, nullptr, nullptr, nullptr, 0, &nray, stream);
  long long *roff = B.get<long long>((size_t)fm->nrcv + 1);
  int *cell = B.get<int>((size_t)nray);
  float *w = B.get<float>((size_t)nray);
  int *ridx = row_index(fm, B);
  if (!roff || !cell || !w || !ridx) return fail("surftime_compute_grad: device allocation");
  if (jdsg_fmst_rays(fm, roff, cell, w, nray, &nray, stream)) return 1;
  // resid = data - obst, in float as the reference forms it (:758)
  {
    std::vector<float> a(fm->nrcv), b(fm->nrcv);
    FCHK(hipMemcpy(a.data(), data, sizeof(float) * fm->nrcv, hipMemcpyDeviceToHost));
    FCHK(hipMemcpy(b.data(), obst, sizeof(float) * fm->nrcv, hipMemcpyDeviceToHost));
    for (int q = 0; q < fm->nrcv; ++q) a[q] = a[q] - b[q];
    FCHK(hipMemcpy(resid, a.data(), sizeof(float) * fm->nrcv, hipMemcpyHostToDevice));
  }
  if (jdsg_frechet_grad(d.nx, d.ny, d.nz, d.kmax, ker, fm->nrcv, ridx, roff, cell, w, resid, vs,
                        grad, stream))
    return fail(std::string("frechet_grad: ") + jdsg_last_error());
  return 0;
}
