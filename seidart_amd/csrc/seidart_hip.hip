// seidart_hip.hip -- MI355X (gfx950) FDTD time-stepping core of SeidarT.
//
// Hand-written HIP kernels for the four reference schemes plus the C ABI
// declared in include/seidart_hip.h.  One source, templated on the
// arithmetic type T (double or float).  The scheme of every kernel is the
// legacy Fortran loop it cites; the numerical conventions (which psi array
// decays with which profile, which coefficient average feeds which stress)
// are kept exactly, including the legacy's shared psi arrays.
//
// Data layout in HBM (all x fastest, padded grid, one plane per shot):
//   fields        nx*nz (2-D) or nx*ny*nz (2.5-D), precision T
//   coefficients  nx*nz per model, shared by every shot on that model
//   psi           slab storage: x-axis memory variables only on the
//                 i < lx / i >= nx-lx columns, z-axis ones only on the
//                 j < lz / j >= nz-lz rows (likewise y in 2.5-D), so the
//                 interior carries no C-PML state at all.
//   traces        [it][comp][rec] per shot
//
// Execution: every time step is one field-update launch (2-D elastic, a
// single pass) or two (stress then velocity, or H then E) over one
// (shot, tile) work list covering every shot of every model, plus one
// receiver-gather launch.  Work items are ordered so that all shots of one
// tile land on the same XCD (blocks are dealt to the 8 XCDs round-robin),
// so a model's coefficient tile is read from HBM once and served from that
// XCD's L2 to the other shots.  The adjoint kernels (2-D elastic and EM)
// sit after the forward kernels of their scheme.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/seidart_hip.h"

#ifndef SDT_NT_STORE
#define SDT_NT_STORE 0
#endif
#ifndef SDT_NT_LOAD
#define SDT_NT_LOAD 0
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(SDT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int TX = 64;  // tile width in x (one wavefront row)
constexpr int TZ = 4;   // tile height in z
constexpr int NT = TX * TZ;

// Device copies of the descriptors.
struct DModel {
  int nx, ny, nz, lx, ly, lz;
  double dx, dy, dz, dt;
  double scal[4];
  const void* coef[40];
  const void* prof[3][6];
  const void* aux;
};

struct DShot {
  int model, isrc, jsrc, ksrc;
  void* f[9];
  void* psi;
  void* g[9];     // second state buffer (ping-pong kinds)
  void* psi2;
  const void* src[3];
  int nrec;
  const int* rec;
  void* trace;
  int* blowup;
  const void* msrc[6];  // stress increments xx yy zz xy xz yz (elastic)
  // receivers by row (ELASTIC2, EM2): those on row k are rxid[rrow[k] .. rrow[k+1]),
  // each a pair (column, receiver index), columns ascending
  const int* rrow;
  const int* rxid;
};

// Adjoint-side buffers of one shot (elastic 2-D), see sdt_adj_shot.
struct DAdj {
  void* trail[2];   // forward states, slot k at trail + k * slot (two buffers)
  long slot;        // elements of one state: fields then psi
  int trail_len;    // slots available
  void* lam[2];     // adjoint state ping-pong: lambda_t in lam[t & 1]
  void* work;       // 12 scratch planes nx*nz
  const void* resid;
  double* grad;     // [12][nz][nx]
  double* gsrc;     // [nt][3]
  float* g32;       // FP32 plans: per-call staging of grad (same layout)
};

// FP32 gradient staging of one distinct grad array, flushed into it.
struct DFlush {
  double* grad;
  float* g32;
  long n;           // cells of the model
};

enum { PK = 0, PA, PB, PKH, PAH, PBH };

// Array pointers live in descriptors in device memory, so the compiler would
// treat them as generic (flat) addresses; flat loads also count against the
// LDS/scalar wait counter, which serialises every shuffle and scalar load
// with all outstanding memory traffic.  Every array access therefore goes
// through explicitly global (address space 1) pointers.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using gcptr = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gcptr<T> cp(const void* p) {
  return (gcptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> mp(void* p) {
  return (gptr<T>)p;
}

// Read-only array through the constant address space: a wave-uniform index
// becomes a scalar load into SGPRs (only for data no kernel of the launch
// writes).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* kp(const void* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

// C-PML profile table of a model: P[axis * 6 + q] (q as in PK..PBH).
template <typename T>
struct Prof {
  const void* const* p;
  __device__ explicit Prof(const void* const* q) : p(q) {}
  __device__ gcptr<T> operator[](int q) const { return cp<T>(p[q]); }
};

// Receiver samples of step itn at row b of one shot (2-D kinds), called by
// every lane of a wavefront that owns lanes lo..hi of row b (lane l holds
// column col0 + l; a, c: the two components on each lane).  The receivers
// of the row (DShot::rrow / rxid) are taken 64 at a time, one per lane,
// which fetches its value from the lane that holds its column and writes it
// if this wavefront owns that column; trace layout [it][comp][rec].
template <typename T>
__device__ __forceinline__ void gather_row(const DShot& S, int b, int itn, int lane, int col0, int lo,
                                           int hi, T a, T c) {
  const int q0 = S.rrow[b], q1 = S.rrow[b + 1];
  const int nrec = S.nrec;
  auto tr = mp<T>(S.trace) + (long)itn * 2 * nrec;
  for (int e0 = q0; e0 < q1; e0 += 64) {
    const int e = e0 + lane;
    const bool has = e < q1;
    const int rx = has ? S.rxid[2 * e] : col0 - 1, id = has ? S.rxid[2 * e + 1] : 0;
    const int src = rx - col0;  // lane holding column rx
    const bool mine = has && src >= lo && src <= hi;
    const T va = __shfl(a, src & 63, 64), vc = __shfl(c, src & 63, 64);
    if (mine) {
      tr[id] = va;
      tr[nrec + id] = vc;
    }
  }
}

// Flag instability: |v|^2 > 1e50 or non-finite (seismicfdtd.f95:378-379).
__device__ __forceinline__ void check_blowup(double v2, int* flag, int it) {
  if (!(v2 <= 1.0e50)) atomicMin(flag, it);
}

// slab offset helpers: position of column i inside the two x slabs
__device__ __forceinline__ int slab_pos(int i, int n, int l) { return i < l ? i : i - n + 2 * l; }
__device__ __forceinline__ bool in_slab(int i, int n, int l) { return i < l || i >= n - l; }

// ===========================================================================
// ELASTIC2: seismic2, legacy/fortran/seismicfdtd.f95:12-385
//
// One fused kernel per time step.  A wavefront owns a tile of EW columns x
// eh rows of one shot and walks it top to bottom.  Lane l holds column
// xo - 1 + l; lanes 1..EW own their column, lane 0 and lanes EW+1, EW+2 are
// halo.  At row r the wave updates the stresses of row r (both legacy stress
// loops, sharing the velocity-derivative psi as the legacy does) and then
// the velocities of row r-1, which need the new stresses of rows r-2..r.
// The previous and next velocity rows live in registers; x neighbours come
// from the neighbouring lanes (wave shuffles).  Stresses of the halo row
// above and below the tile and of the halo lanes are recomputed, so every
// tile is independent; the state therefore ping-pongs between two buffers
// (step `it` reads buffer it&1 and writes the other), which keeps halo reads
// of neighbouring tiles on old values.  Each cell's fields, psi and
// coefficients cross HBM once per step.
//
// psi slabs: x [nz][2lx]: dvx_dx dvz_dx dsxx_dx dsxz_dx (+ dvx_dx dvz_dx of
//                          the shear loop in the split variant)
//            z [2lz][nx]: dvz_dz dvx_dz dszz_dz dsxz_dz (+ dvz_dz dvx_dz)
// ===========================================================================
#ifndef SDT_EL2_ROWS
#define SDT_EL2_ROWS 32
#endif
#ifndef SDT_EL2_GRID
#define SDT_EL2_GRID 1000000
#endif
#ifndef SDT_EL2_WAVES_PER_SIMD
#define SDT_EL2_WAVES_PER_SIMD 4
#endif
#ifndef SDT_EL2_PREFETCH
#define SDT_EL2_PREFETCH 0
#endif
#ifndef SDT_EL2_WAVES_PER_BLOCK
#define SDT_EL2_WAVES_PER_BLOCK 4
#endif
constexpr int EW = 61;            // owned columns per wavefront
// Tile shape: wavefronts per block (side by side in x; a template
// parameter, 1, 2 or 4) and rows per tile (run time); sdt_plan_set_tiles
// changes both, and the autotuner picks them per plan.

// Array addressed by 32-bit element offsets: the scalar base plus a 32-bit
// vector byte offset maps onto the global load/store form with an SGPR base
// and a VGPR offset, so no per-access 64-bit address arithmetic.  Every
// array it indexes is smaller than 4 GiB (checked at plan creation).
template <typename T>
struct A32 {
  using PT = __attribute__((address_space(1))) T*;
  PT p;
  __device__ explicit A32(const void* q) : p((PT)q) {}
  __device__ __forceinline__ __attribute__((address_space(1))) T& operator[](uint32_t e) const {
    return *(PT)((__attribute__((address_space(1))) char*)p + e * (uint32_t)sizeof(T));
  }
};

template <typename T>
__device__ __forceinline__ T lane_next(T v) {  // value of lane l+1
  return __shfl_down(v, 1, 64);
}
template <typename T>
__device__ __forceinline__ T lane_prev(T v) {  // value of lane l-1
  return __shfl_up(v, 1, 64);
}

// The same by DPP wavefront shifts (a VALU move the compiler folds into the
// instruction that uses it, no LDS round trip): dpp_next gives lane l the
// value of lane l+1, dpp_prev that of lane l-1; the lane with no neighbour
// (63, resp. 0) gets 0.  Faster in the two-step kernel; in the one-step
// kernels the LDS form measured faster (they are latency-bound at lower
// occupancy, and a VALU result feeding a DPP read needs wait states).
__device__ __forceinline__ int dpp_next_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_prev_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
template <typename T>
__device__ __forceinline__ T dpp_next(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_next_i(__builtin_bit_cast(int, v)));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_next_i((int)(unsigned)b), hi = (unsigned)dpp_next_i((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}
template <typename T>
__device__ __forceinline__ T dpp_prev(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_prev_i(__builtin_bit_cast(int, v)));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_prev_i((int)(unsigned)b), hi = (unsigned)dpp_prev_i((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

template <typename T, bool TRAIL, bool SPLIT, bool PML, bool R1>
__device__ __forceinline__ void el2_tile(const DModel* __restrict__ models,
                                         const DShot* __restrict__ shots, const int4 wi, const int it,
                                         const int lane, const DAdj* __restrict__ adj, const int slot,
                                         const int tb, const int eh, const int par);

// Work: a block of WB wavefronts takes one item (shot, xo, zo): wave w owns
// columns xo + w*EW .. +EW-1 of rows zo .. zo+eh-1, so the block's row
// accesses are WB*EW contiguous cells of each array.  The grid is
// persistent: block b takes items b, b + gridDim.x, ...; gridDim.x is a
// multiple of 8, so item w always runs on XCD w % 8.
// TRAIL: read the state of step it from trail slot rel (the run starts at
// slot 0) and write the new state into the next slot (forward recomputation
// for the adjoint).
// SPLIT: the split C-PML variant (see SDT_SPLIT_CPML).
// R1: tiles of one row (eh == 1), whose three row steps take their loads
// all at once (a latency-bound shape: register budget of one wavefront per
// SIMD).
// The state of step it is read from buffer it & 1 (A = field/psi, B =
// field_b/psi_b) and written to the other; with prun (the first step of a
// run of two-step launches, k_el2_step2) it is read from buffer
// ((it - *prun) / 2) & 1 instead, the buffer those launches leave it in.
template <typename T, bool TRAIL, bool SPLIT, int WB, bool R1 = false>
__global__ void __launch_bounds__(64 * WB, R1 ? 1 : (sizeof(T) == 4 ? SDT_EL2_WAVES_PER_SIMD : 3))
k_el2_step(const DModel* __restrict__ models, const DShot* __restrict__ shots,
           const int4* __restrict__ work, int nitems, const int* __restrict__ base, int rel,
           const DAdj* __restrict__ adj, const int* __restrict__ tsel, int eh,
           const int* __restrict__ prun) {
  // wave-uniform values kept in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int it = __builtin_amdgcn_readfirstlane(*base + rel);
  const int par = prun ? __builtin_amdgcn_readfirstlane(((it - *prun) >> 1) & 1) : (it & 1);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    // .w: the shot's model, so its record loads beside the shot's
    const int4 wi = make_int4(__builtin_amdgcn_readfirstlane(wv.x),
                              __builtin_amdgcn_readfirstlane(wv.y) + wave * EW,
                              __builtin_amdgcn_readfirstlane(wv.z), __builtin_amdgcn_readfirstlane(wv.w));
    if (wi.x < 0) continue;
    const DModel& M = models[wi.w];
    if (wi.y >= M.nx) continue;
    // a wave whose columns and rows (halo included) avoid every C-PML slab
    // runs the body without memory variables
    const bool interior = wi.y - 1 >= M.lx && wi.y + 62 < M.nx - M.lx && wi.z - 2 >= M.lz &&
                          wi.z + eh + 1 < M.nz - M.lz;
    const int tb = TRAIL ? __builtin_amdgcn_readfirstlane(*tsel) : 0;
    if (interior)
      el2_tile<T, TRAIL, SPLIT, false, R1>(models, shots, wi, it, threadIdx.x & 63, adj, rel, tb, eh, par);
    else
      el2_tile<T, TRAIL, SPLIT, true, R1>(models, shots, wi, it, threadIdx.x & 63, adj, rel, tb, eh, par);
  }
}

template <typename T, bool TRAIL, bool SPLIT, bool PML, bool R1>
__device__ __forceinline__ void el2_tile(const DModel* __restrict__ models,
                                         const DShot* __restrict__ shots, const int4 wi, const int it,
                                         const int lane, const DAdj* __restrict__ adj, const int slot,
                                         const int tb, const int eh, const int par) {
  // Every fused multiply-add is written out; nothing else is contracted, so
  // the interior and C-PML bodies (and every tile shape) round identically.
#pragma clang fp contract(off)
  const DShot& S = shots[wi.x];
  const DModel& M = models[wi.w];  // the shot's model (carried by the work item)
  const int nx = M.nx, nz = M.nz;
  const int xo = wi.y, zo = wi.z;
  const int i = xo - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);  // clamped column: every load stays in bounds
  const bool own = incol && lane >= 1 && lane <= EW;
  const int zend = min(zo + eh, nz);
  const bool has_rec = !TRAIL && S.nrec > 0 && S.rrow[zend] > S.rrow[zo];  // receivers on own rows
  const long nn = (long)nx * nz;
  const T* tin = nullptr;
  T* tout = nullptr;
  if (TRAIL) {
    const DAdj& A = adj[wi.x];
    tin = static_cast<const T*>(A.trail[tb]) + (long)slot * A.slot;
    tout = static_cast<T*>(A.trail[tb]) + (long)(slot + 1) * A.slot;
  }
  const A32<const T> vxi(TRAIL ? tin : (par ? S.g[0] : S.f[0]));
  const A32<const T> vzi(TRAIL ? tin + nn : (par ? S.g[1] : S.f[1]));
  const A32<const T> sxxi(TRAIL ? tin + 2 * nn : (par ? S.g[2] : S.f[2]));
  const A32<const T> szzi(TRAIL ? tin + 3 * nn : (par ? S.g[3] : S.f[3]));
  const A32<const T> sxzi(TRAIL ? tin + 4 * nn : (par ? S.g[4] : S.f[4]));
  const A32<T> vxo(TRAIL ? tout : (par ? S.f[0] : S.g[0]));
  const A32<T> vzo(TRAIL ? tout + nn : (par ? S.f[1] : S.g[1]));
  const A32<T> sxxo(TRAIL ? tout + 2 * nn : (par ? S.f[2] : S.g[2]));
  const A32<T> szzo(TRAIL ? tout + 3 * nn : (par ? S.f[3] : S.g[3]));
  const A32<T> sxzo(TRAIL ? tout + 4 * nn : (par ? S.f[4] : S.g[4]));
  const A32<const T> pin(TRAIL ? tin + 5 * nn : (par ? S.psi2 : S.psi));
  const A32<T> pout(TRAIL ? tout + 5 * nn : (par ? S.psi : S.psi2));
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dt = T(M.dt);
  const int lx = M.lx, lz = M.lz;
  const uint32_t SX = (uint32_t)nz * 2 * lx, SZ = (uint32_t)2 * lz * nx;
  const uint32_t ZB = (SPLIT ? 6 : 4) * SX;  // first z-slab array
  const bool inx = PML && incol && in_slab(i, nx, lx);
  const int xs = inx ? slab_pos(i, nx, lx) : 0;
  const Prof<T> P(&M.prof[0][0]);
  const T ikx = P[PK][ic], ax = P[PA][ic], bx = P[PB][ic];
  const T ikxh = P[PKH][ic], axh = P[PAH][ic], bxh = P[PBH][ic];
  const auto ikz = P[12 + PK], az = P[12 + PA], bz = P[12 + PB];
  const auto ikzh = P[12 + PKH], azh = P[12 + PAH], bzh = P[12 + PBH];
  const auto brow = cp<T>(M.aux);
  const A32<const T> N11(M.coef[0]), N13(M.coef[1]), N15(M.coef[2]);
  const A32<const T> N33(M.coef[3]), N35(M.coef[4]);
  const A32<const T> S15(M.coef[5]), S35(M.coef[6]), S55(M.coef[7]);
  const A32<const T> DTX(M.coef[8]), DTZ(M.coef[9]);
  const A32<const T> OMX(M.coef[10]), OMZ(M.coef[11]);
  const bool mono = M.coef[2] != nullptr;   // c15 / c35 terms present
  const bool atten = M.coef[10] != nullptr; // gamma present
  const bool can_n = incol && lane <= EW + 1 && i <= nx - 2;          // normal stress (needs i+1)
  const bool can_s = incol && lane >= 1 && lane <= EW + 1 && i >= 1;  // shear (needs i-1)
  const bool ux = own && i >= 1, uz = own && i <= nx - 2;              // vx / vz columns
  const bool is_src_col = (i == S.isrc);
  const int ksrc = S.ksrc;
  const auto srcx = cp<T>(S.src[0]), srcz = cp<T>(S.src[2]);
  const T incx = (srcx != nullptr) ? srcx[it] : T(0);
  const T incz = (srcz != nullptr) ? srcz[it] : T(0);
  // stress-rate (moment-tensor) source: xx, zz, xz increments of this step
  const bool msrc = S.msrc[0] || S.msrc[2] || S.msrc[4];
  const T incsxx = S.msrc[0] ? cp<T>(S.msrc[0])[it] : T(0);
  const T incszz = S.msrc[2] ? cp<T>(S.msrc[2])[it] : T(0);
  const T incsxz = S.msrc[4] ? cp<T>(S.msrc[4])[it] : T(0);

  auto row = [&](int r) { return (uint32_t)(r < 0 ? 0 : (r >= nz ? nz - 1 : r)) * nx + ic; };
  auto valid = [&](int r) { return incol && r >= 0 && r < nz; };
  T vx_m = valid(zo - 2) ? vxi[row(zo - 2)] : T(0);  // velocity row r-1
  T vz_m = valid(zo - 2) ? vzi[row(zo - 2)] : T(0);
  T vx_c = valid(zo - 1) ? vxi[row(zo - 1)] : T(0);  // velocity row r
  T vz_c = valid(zo - 1) ? vzi[row(zo - 1)] : T(0);
  T sxx_p = 0, szz_p = 0, sxz_p = 0, sxz_pp = 0;      // new stresses of rows r-1, r-2

  // Every load a row needs (addresses clamped in bounds).  With
  // SDT_EL2_PREFETCH the loads of row r+1 are issued before row r is
  // computed, so each wave keeps two rows of loads in flight.
  struct Row {
    T vxn, vzn, sxx, szz, sxz, n11, n13, n33, s55, n15, n35, s15, s35, dtx, dtz, omx, omz;
    T m_vxx, m_vzx, m_vzz, m_vxz, s_vxx, s_vzx, s_vzz, s_vxz, m_sxx, m_sxzx, m_szz, m_sxzz;
  };
  auto load_row = [&](int r) {
    Row L;
    const int rr = r - 1;
    const uint32_t c = row(r), cv = row(rr), cn = row(r + 1);
    L.vxn = vxi[cn];
    L.vzn = vzi[cn];
    L.sxx = sxxi[c];
    L.szz = szzi[c];
    L.sxz = sxzi[c];
    L.n11 = N11[c];
    L.n13 = N13[c];
    L.n33 = N33[c];
    L.s55 = S55[c];
    // monoclinic couplings and attenuation are optional (null = 0 and 1)
    L.n15 = mono ? N15[c] : T(0);
    L.n35 = mono ? N35[c] : T(0);
    L.s15 = mono ? S15[c] : T(0);
    L.s35 = mono ? S35[c] : T(0);
    L.dtx = DTX[cv];
    L.dtz = DTZ[cv];
    L.omx = atten ? OMX[cv] : T(1);
    L.omz = atten ? OMZ[cv] : T(1);
    const int rs = r < 0 ? 0 : (r >= nz ? nz - 1 : r);
    const int rvs = rr < 0 ? 0 : (rr >= nz ? nz - 1 : rr);
    const bool inz = PML && in_slab(rs, nz, lz), inzv = PML && in_slab(rvs, nz, lz);
    const uint32_t oxs = inx ? (uint32_t)rs * 2 * lx + xs : 0, oxvs = inx ? (uint32_t)rvs * 2 * lx + xs : 0;
    const uint32_t oz = ZB + (inz ? (uint32_t)slab_pos(rs, nz, lz) * nx + ic : 0);
    const uint32_t ozv = ZB + (inzv ? (uint32_t)slab_pos(rvs, nz, lz) * nx + ic : 0);
    // psi loads only where the cell has that memory (a zero-width slab has no storage)
    const bool inzc = incol && inz, inzvc = incol && inzv;
    L.m_vxx = inx ? pin[oxs] : T(0);  // x-slab, stress row
    L.m_vzx = inx ? pin[SX + oxs] : T(0);
    L.m_vzz = inzc ? pin[oz] : T(0);  // z-slab, stress row
    L.m_vxz = inzc ? pin[SZ + oz] : T(0);
    // split C-PML: the shear loop's own memory (x slab arrays 4, 5; z 4, 5)
    L.s_vxx = (SPLIT && inx) ? pin[4 * SX + oxs] : T(0);
    L.s_vzx = (SPLIT && inx) ? pin[5 * SX + oxs] : T(0);
    L.s_vzz = (SPLIT && inzc) ? pin[4 * SZ + oz] : T(0);
    L.s_vxz = (SPLIT && inzc) ? pin[5 * SZ + oz] : T(0);
    L.m_sxx = inx ? pin[2 * SX + oxvs] : T(0);  // velocity row
    L.m_sxzx = inx ? pin[3 * SX + oxvs] : T(0);
    L.m_szz = inzvc ? pin[2 * SZ + ozv] : T(0);
    L.m_sxzz = inzvc ? pin[3 * SZ + ozv] : T(0);
    return L;
  };
  // (FP64 keeps one row in flight: two would spill at its register budget)
  constexpr bool PF = SDT_EL2_PREFETCH && sizeof(T) == 4;
  Row ahead{};
  if constexpr (PF) ahead = load_row(zo - 1);

  // row step r: the stresses of row r, then the velocities of row r - 1
  auto step_row = [&](const int r, const Row& L) {
    const int rr = r - 1;                             // velocity row
    const uint32_t c = row(r), cv = row(rr);
    const bool rv = valid(r);
    const T ld_vxn = L.vxn, ld_vzn = L.vzn, ld_sxx = L.sxx, ld_szz = L.szz, ld_sxz = L.sxz;
    const T n11 = L.n11, n13 = L.n13, n33 = L.n33, s55 = L.s55;
    const T n15 = L.n15, n35 = L.n35, s15 = L.s15, s35 = L.s35;
    const T dtx = L.dtx, dtz = L.dtz, omx = L.omx, omz = L.omz;
    const int rs = r < 0 ? 0 : (r >= nz ? nz - 1 : r);
    const int rvs = rr < 0 ? 0 : (rr >= nz ? nz - 1 : rr);
    const bool inz = PML && in_slab(rs, nz, lz), inzv = PML && in_slab(rvs, nz, lz);
    const uint32_t ox = (uint32_t)rs * 2 * lx + xs, oxv = (uint32_t)rvs * 2 * lx + xs;
    const uint32_t oz = ZB + (inz ? (uint32_t)slab_pos(rs, nz, lz) * nx + ic : 0);
    const uint32_t ozv = ZB + (inzv ? (uint32_t)slab_pos(rvs, nz, lz) * nx + ic : 0);
    T m_vxx = L.m_vxx, m_vzx = L.m_vzx, m_vzz = L.m_vzz, m_vxz = L.m_vxz;
    T s_vxx = L.s_vxx, s_vzx = L.s_vzx, s_vzz = L.s_vzz, s_vxz = L.s_vxz;
    T m_sxx = L.m_sxx, m_sxzx = L.m_sxzx, m_szz = L.m_szz, m_sxzz = L.m_sxzz;
    const T bzr = bz[rs], azr = az[rs], ikzr = ikz[rs], bzhr = bzh[rs], azhr = azh[rs], ikzhr = ikzh[rs];
    const T browr = SPLIT ? bxh : brow[rs];  // b_x_half(j) (A2), or b_x_half(i)
    const T bzv = bz[rvs], azv = az[rvs], ikzv = ikz[rvs], bzhv = bzh[rvs], azhv = azh[rvs], ikzhv = ikzh[rvs];
    const bool nxt = (r + 1 <= zend) && valid(r + 1);
    const T vx_n = nxt ? ld_vxn : T(0), vz_n = nxt ? ld_vzn : T(0);  // velocity row r+1

    // ---- stresses of row r (seismicfdtd.f95:247-312) ----
    const T vx_e = lane_next(vx_c), vz_e = lane_next(vz_c);
    const T vx_w = lane_prev(vx_c), vz_w = lane_prev(vz_c);
    T sxx_c = rv ? ld_sxx : T(0), szz_c = rv ? ld_szz : T(0), sxz_c = rv ? ld_sxz : T(0);
    const bool inx_r = inx && rv, inz_r = incol && rv && inz;
    if (rv && can_n && r >= 1) {
      T dvx_dx = (vx_e - vx_c) * idx;
      T dvz_dz = (vz_c - vz_m) * idz;
      T dvz_dx = (vz_e - vz_c) * idx;
      T dvx_dz = (vx_c - vx_m) * idz;
      if (inx_r) {
        m_vxx = fma(browr, m_vxx, axh * dvx_dx);  // :255
        m_vzx = fma(bx, m_vzx, ax * dvz_dx);
        dvx_dx = fma(dvx_dx, ikxh, m_vxx);
        dvz_dx = fma(dvz_dx, ikx, m_vzx);
      }
      if (inz_r) {
        m_vzz = fma(bzr, m_vzz, azr * dvz_dz);
        m_vxz = fma(bzhr, m_vxz, azhr * dvx_dz);
        dvz_dz = fma(dvz_dz, ikzr, m_vzz);
        dvx_dz = fma(dvx_dz, ikzhr, m_vxz);
      }
      const T sh = dvz_dx + dvx_dz;
      sxx_c = fma(fma(n15, sh, fma(n13, dvz_dz, n11 * dvx_dx)), dt, sxx_c);
      szz_c = fma(fma(n35, sh, fma(n33, dvz_dz, n13 * dvx_dx)), dt, szz_c);
    }
    if (rv && can_s && r <= nz - 2 && r < zend) {  // same psi arrays (A1) unless split
      T& q_vxx = SPLIT ? s_vxx : m_vxx;
      T& q_vzx = SPLIT ? s_vzx : m_vzx;
      T& q_vzz = SPLIT ? s_vzz : m_vzz;
      T& q_vxz = SPLIT ? s_vxz : m_vxz;
      T dvx_dx = (vx_c - vx_w) * idx;
      T dvz_dz = (vz_n - vz_c) * idz;
      T dvz_dx = (vz_c - vz_w) * idx;
      T dvx_dz = (vx_n - vx_c) * idz;
      if (inx_r) {
        q_vxx = fma(bxh, q_vxx, axh * dvx_dx);
        q_vzx = fma(bx, q_vzx, ax * dvz_dx);
        dvx_dx = fma(dvx_dx, ikxh, q_vxx);
        dvz_dx = fma(dvz_dx, ikx, q_vzx);
      }
      if (inz_r) {
        q_vzz = fma(bzr, q_vzz, azr * dvz_dz);
        q_vxz = fma(bzhr, q_vxz, azhr * dvx_dz);
        dvz_dz = fma(dvz_dz, ikzr, q_vzz);
        dvx_dz = fma(dvx_dz, ikzhr, q_vxz);
      }
      sxz_c = fma(fma(s55, dvz_dx + dvx_dz, fma(s35, dvz_dz, s15 * dvx_dx)), dt, sxz_c);
    }
    if (msrc && rv && is_src_col && r == ksrc) {  // every lane that computes the cell adds it
      sxx_c += incsxx;
      szz_c += incszz;
      sxz_c += incsxz;
    }
    if (own && rv && r >= zo && r < zend) {
      sxxo[c] = sxx_c;
      szzo[c] = szz_c;
      sxzo[c] = sxz_c;
      if (inx) {
        pout[ox] = m_vxx;
        pout[SX + ox] = m_vzx;
        if (SPLIT) {
          pout[4 * SX + ox] = s_vxx;
          pout[5 * SX + ox] = s_vzx;
        }
      }
      if (inz) {
        pout[oz] = m_vzz;
        pout[SZ + oz] = m_vxz;
        if (SPLIT) {
          pout[4 * SZ + oz] = s_vzz;
          pout[5 * SZ + oz] = s_vxz;
        }
      }
    }

    // ---- velocities of row rr = r - 1 (seismicfdtd.f95:317-375) ----
    const T sxx_w = lane_prev(sxx_p);  // sigma_xx(i-1, rr)
    const T sxz_e = lane_next(sxz_p);  // sigma_xz(i+1, rr)
    T gvx = 0, gvz = 0;  // new velocities of the owned cells (receiver samples)
    if (own && rr >= zo && rr < zend) {
      T vxn = vx_m, vzn = vz_m;
      if (ux && rr >= 1) {
        T a = (sxx_p - sxx_w) * idx;
        T b = (sxz_p - sxz_pp) * idz;
        if (inx) {
          m_sxx = fma(bx, m_sxx, ax * a);
          a = fma(a, ikx, m_sxx);
        }
        if (inzv) {
          m_sxzz = fma(bzv, m_sxzz, azv * b);
          b = fma(b, ikzv, m_sxzz);
        }
        vxn = fma(vxn, omx, (a + b) * dtx);
      }
      if (uz && rr <= nz - 2) {
        T a = (sxz_e - sxz_p) * idx;
        T b = (szz_c - szz_p) * idz;
        if (inx) {
          m_sxzx = fma(bxh, m_sxzx, axh * a);
          a = fma(a, ikxh, m_sxzx);
        }
        if (inzv) {
          m_szz = fma(bzhv, m_szz, azhv * b);
          b = fma(b, ikzhv, m_szz);
        }
        vzn = fma(vzn, omz, (a + b) * dtz);
      }
      if (inx) { pout[2 * SX + oxv] = m_sxx; pout[3 * SX + oxv] = m_sxzx; }
      if (inzv) { pout[2 * SZ + ozv] = m_szz; pout[3 * SZ + ozv] = m_sxzz; }
      if (is_src_col && rr == ksrc) { vxn += incx; vzn += incz; }   // source, :359-361
      if (i == 0 || i == nx - 1 || rr == 0 || rr == nz - 1) { vxn = 0; vzn = 0; }  // :365-375
      check_blowup((double)vxn * vxn + (double)vzn * vzn, S.blowup, it);
      vxo[cv] = vxn;
      vzo[cv] = vzn;
      gvx = vxn;
      gvz = vzn;
    }
    if (has_rec && rr >= zo && rr < zend) gather_row<T>(S, rr, it, lane, xo - 1, 1, EW, gvx, gvz);
    sxz_pp = sxz_p;
    sxx_p = sxx_c;
    szz_p = szz_c;
    sxz_p = sxz_c;
    vx_m = vx_c;
    vz_m = vz_c;
    vx_c = vx_n;
    vz_c = vz_n;
  };
  if constexpr (R1) {
    // one row (zend == zo + 1): the loads of rows zo-1, zo, zo+1 read only
    // the previous state, so all three are in flight before the first step
    const Row L0 = load_row(zo - 1), L1 = load_row(zo), L2 = load_row(zo + 1);
    step_row(zo - 1, L0);
    step_row(zo, L1);
    step_row(zo + 1, L2);
  } else {
    for (int r = zo - 1; r <= zend; ++r) {
      Row L;
      if constexpr (PF) {
        L = ahead;
        ahead = load_row(r + 1);
      } else {
        L = load_row(r);
      }
      step_row(r, L);
    }
  }
}

// ===========================================================================
// ELASTIC2, two time steps per launch.
//
// The walk of k_el2_step with the second step pipelined behind the first.
// At row r a wavefront computes
//   A  the stresses of step it at row r      (velocities of it-1, rows r-1..r+1)
//   B  the velocities of step it at row r-1  (stresses from A, rows r-2..r)
//   C  the stresses of step it+1 at row r-2  (velocities from B, rows r-3..r-1)
//   D  the velocities of step it+1 at row r-3 (stresses from C, rows r-4..r-2)
// so the state of step it lives only in registers and each field, memory
// variable and coefficient crosses HBM once per two steps.
// One step makes a cell depend on columns i-1 .. i+2 and rows k-2 .. k+1 of
// the previous state (the shear stress updates the C-PML memory the normal
// stresses just updated, quirk A1, so it sees the normal loop's neighbours);
// two steps on columns i-2 .. i+4 and rows k-4 .. k+2.  Lane l therefore
// holds column xo - 2 + l and lanes 2 .. 2+EW2-1 are exact after two steps;
// rows zo-4 .. zo+eh+1 are read to produce rows zo .. zo+eh-1.  Every stage
// evaluates the expressions of k_el2_step in the same order, so both kernels
// give identical bits.  The state of step it is never stored, so this kernel
// also writes the receiver samples of both steps.
// The state of step it is read from buffer ((it - *it0) / 2) & 1 and the
// state of it+2 written to the other (it0: first step of the run).
// ===========================================================================
// Owned columns per wavefront (at most 58: two halo lanes on the left, four
// on the right).  56 makes every wavefront's row of stores start and end on
// a 32-byte boundary (56 floats = 224 bytes, tiles at multiples of 56), so
// no HBM sector is written in pieces by two wavefronts: 134.3 against 131.4
// Gcell-updates/s with 58 on the 96-shot layered workload, same bits.
#ifndef SDT_EL2X2_EW
#define SDT_EL2X2_EW 56
#endif
constexpr int EW2 = SDT_EL2X2_EW;
static_assert(EW2 >= 1 && EW2 <= 58, "two-step wavefront: 2 halo lanes left, 4 right");
#ifndef SDT_EL2X2_WAVES
#define SDT_EL2X2_WAVES 4   // FP32 wavefronts per SIMD the register budget allows
#endif
#ifndef SDT_EL2X2_MIN_CELLS
#define SDT_EL2X2_MIN_CELLS 4000000LL  // cells per step from which plans default to two steps per launch
#endif
#ifndef SDT_EL2X2_PREFETCH
#define SDT_EL2X2_PREFETCH 0
#endif
#ifndef SDT_EL2X2_WAVES_SPLIT
// the same for the split C-PML variant's edge kernel: at 4 wavefronts it
// spills a few registers and is still faster than at 3 (1.178 against
// 1.228 ms per two-step launch pair of the 96-shot layered workload)
#define SDT_EL2X2_WAVES_SPLIT 4
#endif
#ifndef SDT_EL2X2_WAVES_INT
#define SDT_EL2X2_WAVES_INT 5  // the same for the interior-only kernel
#endif

template <typename T, bool SPLIT, int BODY>
__device__ __forceinline__ void el2x2_tile(const DModel* __restrict__ models,
                                           const DShot* __restrict__ shots, const int4 wi, const int it,
                                           const int q, const int lane, const int eh) {
#pragma clang fp contract(off)
  // BODY 0: an interior wave (no C-PML) reads only columns 1..nx-2 and rows
  // 1..nz-2, so every edge condition of the scheme holds on every lane;
  // BODY 1: its columns are interior (every condition on the column holds,
  // no x slab) but its rows may reach a z slab or the outer rows;
  // BODY 2: anything else; BODY 3 (below): x-slab columns of a tile whose
  // rows are all off the z slabs and the outer rows.
  // BODY 3: as 2, but every row the wave reads is off the z slabs and the
  // outer rows (x-slab columns of a z-interior tile).
  constexpr bool INT = BODY == 0, XI = BODY < 2, PML = BODY > 0, ZI = BODY == 3;
  const DShot& S = shots[wi.x];
  const DModel& M = models[S.model];
  const int nx = M.nx, nz = M.nz;
  const int xo = wi.y, zo = wi.z;
  const int i = xo - 2 + lane;
  const bool incol = XI || (i >= 0 && i < nx);
  const int ic = XI ? i : (incol ? i : (i < 0 ? 0 : nx - 1));  // clamped column: every load stays in bounds
  const bool own = incol && lane >= 2 && lane < 2 + EW2;
  const int ze = min(zo + eh, nz);
  const A32<const T> vxi(q ? S.g[0] : S.f[0]);
  const A32<const T> vzi(q ? S.g[1] : S.f[1]);
  const A32<const T> sxxi(q ? S.g[2] : S.f[2]);
  const A32<const T> szzi(q ? S.g[3] : S.f[3]);
  const A32<const T> sxzi(q ? S.g[4] : S.f[4]);
  const A32<T> vxo(q ? S.f[0] : S.g[0]);
  const A32<T> vzo(q ? S.f[1] : S.g[1]);
  const A32<T> sxxo(q ? S.f[2] : S.g[2]);
  const A32<T> szzo(q ? S.f[3] : S.g[3]);
  const A32<T> sxzo(q ? S.f[4] : S.g[4]);
  const A32<const T> pin(q ? S.psi2 : S.psi);
  const A32<T> pout(q ? S.psi : S.psi2);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dt = T(M.dt);
  const int lx = M.lx, lz = M.lz;
  const uint32_t SX = (uint32_t)nz * 2 * lx, SZ = (uint32_t)2 * lz * nx;
  const uint32_t ZB = (SPLIT ? 6 : 4) * SX;  // first z-slab array
  const bool inx = !XI && incol && in_slab(i, nx, lx);
  const int xs = inx ? slab_pos(i, nx, lx) : 0;
  const Prof<T> P(&M.prof[0][0]);
  const T ikx = P[PK][ic], ax = P[PA][ic], bx = P[PB][ic];
  const T ikxh = P[PKH][ic], axh = P[PAH][ic], bxh = P[PBH][ic];
  // z profiles and b_x_half by row: one value per row, read through the
  // scalar (constant) path into SGPRs; nothing writes them during a launch
  const auto ikz = kp<T>(M.prof[2][PK]), az = kp<T>(M.prof[2][PA]), bz = kp<T>(M.prof[2][PB]);
  const auto ikzh = kp<T>(M.prof[2][PKH]), azh = kp<T>(M.prof[2][PAH]), bzh = kp<T>(M.prof[2][PBH]);
  const auto brow = kp<T>(M.aux);
  const A32<const T> N11(M.coef[0]), N13(M.coef[1]), N15(M.coef[2]);
  const A32<const T> N33(M.coef[3]), N35(M.coef[4]);
  const A32<const T> S15(M.coef[5]), S35(M.coef[6]), S55(M.coef[7]);
  const A32<const T> DTX(M.coef[8]), DTZ(M.coef[9]);
  const A32<const T> OMX(M.coef[10]), OMZ(M.coef[11]);
  const bool mono = M.coef[2] != nullptr;   // c15 / c35 terms present
  const bool atten = M.coef[10] != nullptr; // gamma present
  const bool is_src_col = (i == S.isrc);
  const int ksrc = S.ksrc;
  const auto srcx = cp<T>(S.src[0]), srcz = cp<T>(S.src[2]);
  const T incx0 = srcx ? srcx[it] : T(0), incx1 = srcx ? srcx[it + 1] : T(0);
  const T incz0 = srcz ? srcz[it] : T(0), incz1 = srcz ? srcz[it + 1] : T(0);
  const bool msrc = S.msrc[0] || S.msrc[2] || S.msrc[4];
  const T ms0xx = S.msrc[0] ? cp<T>(S.msrc[0])[it] : T(0), ms1xx = S.msrc[0] ? cp<T>(S.msrc[0])[it + 1] : T(0);
  const T ms0zz = S.msrc[2] ? cp<T>(S.msrc[2])[it] : T(0), ms1zz = S.msrc[2] ? cp<T>(S.msrc[2])[it + 1] : T(0);
  const T ms0xz = S.msrc[4] ? cp<T>(S.msrc[4])[it] : T(0), ms1xz = S.msrc[4] ? cp<T>(S.msrc[4])[it + 1] : T(0);

  auto crow = [&](int r) { return r < 0 ? 0 : (r >= nz ? nz - 1 : r); };
  auto cell = [&](int r) { return (uint32_t)crow(r) * nx + ic; };
  auto valid = [&](int r) { return INT || (incol && r >= 0 && r < nz); };

  struct CN { T n11, n13, n33, s55, n15, n35, s15, s35; };  // stress coefficients of a row
  struct CV { T dtx, dtz, omx, omz; };                      // velocity coefficients of a row
  struct MV { T vxx, vzx, vzz, vxz, svxx, svzx, svzz, svxz; };  // memory of velocity derivatives
  struct MS { T sxx, sxzx, szz, sxzz; };                    // memory of stress derivatives

  auto load_cn = [&](int r) {
    CN k;
    const uint32_t c = cell(r);
    k.n11 = N11[c];
    k.n13 = N13[c];
    k.n33 = N33[c];
    k.s55 = S55[c];
    k.n15 = mono ? N15[c] : T(0);
    k.n35 = mono ? N35[c] : T(0);
    k.s15 = mono ? S15[c] : T(0);
    k.s35 = mono ? S35[c] : T(0);
    return k;
  };
  auto load_cv = [&](int r) {
    CV k;
    const uint32_t c = cell(r);
    k.dtx = DTX[c];
    k.dtz = DTZ[c];
    k.omx = atten ? OMX[c] : T(1);
    k.omz = atten ? OMZ[c] : T(1);
    return k;
  };
  // slab offsets of row r: x slab (valid where inx), z slab (ZB-based)
  auto offx = [&](int r) { return inx ? (uint32_t)crow(r) * 2 * lx + xs : 0u; };
  auto inzs = [&](int r) { return PML && !ZI && in_slab(crow(r), nz, lz); };
  auto offz = [&](int r) { return ZB + (inzs(r) ? (uint32_t)slab_pos(crow(r), nz, lz) * nx + ic : 0u); };
  auto load_mv = [&](int r) {
    MV m{};
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      const bool inzc = incol && inzs(r);
      m.vxx = inx ? pin[ox] : T(0);
      m.vzx = inx ? pin[SX + ox] : T(0);
      m.vzz = inzc ? pin[oz] : T(0);
      m.vxz = inzc ? pin[SZ + oz] : T(0);
      if (SPLIT) {
        m.svxx = inx ? pin[4 * SX + ox] : T(0);
        m.svzx = inx ? pin[5 * SX + ox] : T(0);
        m.svzz = inzc ? pin[4 * SZ + oz] : T(0);
        m.svxz = inzc ? pin[5 * SZ + oz] : T(0);
      }
    }
    return m;
  };
  auto load_ms = [&](int r) {
    MS m{};
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      const bool inzc = incol && inzs(r);
      m.sxx = inx ? pin[2 * SX + ox] : T(0);
      m.sxzx = inx ? pin[3 * SX + ox] : T(0);
      m.szz = inzc ? pin[2 * SZ + oz] : T(0);
      m.sxzz = inzc ? pin[3 * SZ + oz] : T(0);
    }
    return m;
  };
  auto store_mv = [&](int r, const MV& m) {  // own cells only
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      if (inx) {
        pout[ox] = m.vxx;
        pout[SX + ox] = m.vzx;
        if (SPLIT) {
          pout[4 * SX + ox] = m.svxx;
          pout[5 * SX + ox] = m.svzx;
        }
      }
      if (inzs(r)) {
        pout[oz] = m.vzz;
        pout[SZ + oz] = m.vxz;
        if (SPLIT) {
          pout[4 * SZ + oz] = m.svzz;
          pout[5 * SZ + oz] = m.svxz;
        }
      }
    }
  };
  auto store_ms = [&](int r, const MS& m) {  // own cells only
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      if (inx) { pout[2 * SX + ox] = m.sxx; pout[3 * SX + ox] = m.sxzx; }
      if (inzs(r)) { pout[2 * SZ + oz] = m.szz; pout[3 * SZ + oz] = m.sxzz; }
    }
  };

  // Stresses of one step at row a (seismicfdtd.f95:247-312): velocities of
  // rows a-1, a, a+1 in, the row's stresses and memory updated in place.
  auto stress = [&](int a, T vxm, T vzm, T vxc, T vzc, T vxn, T vzn, T& sxx, T& szz, T& sxz,
                    const CN& k, MV& m, T isxx, T iszz, T isxz) {
    const T vx_e = dpp_next(vxc), vz_e = dpp_next(vzc);
    const T vx_w = dpp_prev(vxc), vz_w = dpp_prev(vzc);
    const bool rv = valid(a);
    const int rs = crow(a);
    const bool inx_r = inx && rv, inz_r = incol && rv && inzs(a);
    T bzr = 0, azr = 0, ikzr = 0, bzhr = 0, azhr = 0, ikzhr = 0, browr = 0;
    if constexpr (PML) {
      bzr = bz[rs]; azr = az[rs]; ikzr = ikz[rs];
      bzhr = bzh[rs]; azhr = azh[rs]; ikzhr = ikzh[rs];
      browr = SPLIT ? bxh : brow[rs];  // b_x_half(j) (A2), or b_x_half(i)
    }
    if (INT || (rv && (XI || i <= nx - 2) && a >= 1)) {
      T dvx_dx = (vx_e - vxc) * idx;
      T dvz_dz = (vzc - vzm) * idz;
      T dvz_dx = (vz_e - vzc) * idx;
      T dvx_dz = (vxc - vxm) * idz;
      if (PML && inx_r) {
        m.vxx = fma(browr, m.vxx, axh * dvx_dx);
        m.vzx = fma(bx, m.vzx, ax * dvz_dx);
        dvx_dx = fma(dvx_dx, ikxh, m.vxx);
        dvz_dx = fma(dvz_dx, ikx, m.vzx);
      }
      if (PML && inz_r) {
        m.vzz = fma(bzr, m.vzz, azr * dvz_dz);
        m.vxz = fma(bzhr, m.vxz, azhr * dvx_dz);
        dvz_dz = fma(dvz_dz, ikzr, m.vzz);
        dvx_dz = fma(dvx_dz, ikzhr, m.vxz);
      }
      const T sh = dvz_dx + dvx_dz;
      sxx = fma(fma(k.n15, sh, fma(k.n13, dvz_dz, k.n11 * dvx_dx)), dt, sxx);
      szz = fma(fma(k.n35, sh, fma(k.n33, dvz_dz, k.n13 * dvx_dx)), dt, szz);
    }
    if (INT || (rv && (XI || i >= 1) && a <= nz - 2)) {  // same memory arrays (A1) unless split
      T& q_vxx = SPLIT ? m.svxx : m.vxx;
      T& q_vzx = SPLIT ? m.svzx : m.vzx;
      T& q_vzz = SPLIT ? m.svzz : m.vzz;
      T& q_vxz = SPLIT ? m.svxz : m.vxz;
      T dvx_dx = (vxc - vx_w) * idx;
      T dvz_dz = (vzn - vzc) * idz;
      T dvz_dx = (vzc - vz_w) * idx;
      T dvx_dz = (vxn - vxc) * idz;
      if (PML && inx_r) {
        q_vxx = fma(bxh, q_vxx, axh * dvx_dx);
        q_vzx = fma(bx, q_vzx, ax * dvz_dx);
        dvx_dx = fma(dvx_dx, ikxh, q_vxx);
        dvz_dx = fma(dvz_dx, ikx, q_vzx);
      }
      if (PML && inz_r) {
        q_vzz = fma(bzr, q_vzz, azr * dvz_dz);
        q_vxz = fma(bzhr, q_vxz, azhr * dvx_dz);
        dvz_dz = fma(dvz_dz, ikzr, q_vzz);
        dvx_dz = fma(dvx_dz, ikzhr, q_vxz);
      }
      sxz = fma(fma(k.s55, dvz_dx + dvx_dz, fma(k.s35, dvz_dz, k.s15 * dvx_dx)), dt, sxz);
    }
    if (msrc && a == ksrc && rv && is_src_col) {
      sxx += isxx;
      szz += iszz;
      sxz += isxz;
    }
  };
  // Velocities of one step at row b (seismicfdtd.f95:317-375): the new
  // stresses of rows b-1 (sxz), b, b+1 (szz) in.
  auto vel = [&](int b, T sxx_b, T szz_b, T sxz_b, T sxz_bm, T szz_bn, T vx_old, T vz_old,
                 const CV& k, MS& m, T ix, T iz, T& vxn, T& vzn) {
    const T sxx_w = dpp_prev(sxx_b);  // sigma_xx(i-1, b)
    const T sxz_e = dpp_next(sxz_b);  // sigma_xz(i+1, b)
    const bool rv = valid(b);
    const int rs = crow(b);
    const bool inzv = PML && inzs(b);
    T bzv = 0, azv = 0, ikzv = 0, bzhv = 0, azhv = 0, ikzhv = 0;
    if constexpr (PML) {
      bzv = bz[rs]; azv = az[rs]; ikzv = ikz[rs];
      bzhv = bzh[rs]; azhv = azh[rs]; ikzhv = ikzh[rs];
    }
    vxn = vx_old;
    vzn = vz_old;
    if (INT || (rv && (XI || i >= 1) && b >= 1)) {
      T a = (sxx_b - sxx_w) * idx;
      T c = (sxz_b - sxz_bm) * idz;
      if (PML && inx) {
        m.sxx = fma(bx, m.sxx, ax * a);
        a = fma(a, ikx, m.sxx);
      }
      if (PML && inzv) {
        m.sxzz = fma(bzv, m.sxzz, azv * c);
        c = fma(c, ikzv, m.sxzz);
      }
      vxn = fma(vxn, k.omx, (a + c) * k.dtx);
    }
    if (INT || (rv && (XI || i <= nx - 2) && b <= nz - 2)) {
      T a = (sxz_e - sxz_b) * idx;
      T c = (szz_bn - szz_b) * idz;
      if (PML && inx) {
        m.sxzx = fma(bxh, m.sxzx, axh * a);
        a = fma(a, ikxh, m.sxzx);
      }
      if (PML && inzv) {
        m.szz = fma(bzhv, m.szz, azhv * c);
        c = fma(c, ikzhv, m.szz);
      }
      vzn = fma(vzn, k.omz, (a + c) * k.dtz);
    }
    if (b == ksrc && is_src_col) { vxn += ix; vzn += iz; }   // source, :359-361
    if (!INT && (!rv || (!XI && (i == 0 || i == nx - 1)) || b == 0 || b == nz - 1)) {  // :365-375
      vxn = 0;
      vzn = 0;
    }
  };
  // receiver samples of step itn at row b (an own row; every lane calls)
  const int nrec = S.nrec;
  auto gather = [&](int b, int itn, T vx, T vz) { gather_row<T>(S, b, itn, lane, xo - 2, 2, 1 + EW2, vx, vz); };

  T vx0_m = valid(zo - 4) ? vxi[cell(zo - 4)] : T(0);  // velocities of step it-1, row r-1
  T vz0_m = valid(zo - 4) ? vzi[cell(zo - 4)] : T(0);
  T vx0_c = valid(zo - 3) ? vxi[cell(zo - 3)] : T(0);  // row r
  T vz0_c = valid(zo - 3) ? vzi[cell(zo - 3)] : T(0);
  T s1xx_p = 0, s1zz_p = 0, s1xz_p = 0, s1xx_pp = 0, s1zz_pp = 0, s1xz_pp = 0;  // step it, rows r-1, r-2
  T v1x_m = 0, v1z_m = 0, v1x_c = 0, v1z_c = 0;  // step it, rows r-3, r-2
  T s2xx_p = 0, s2zz_p = 0, s2xz_p = 0, s2xz_pp = 0;  // step it+1, rows r-3 (and r-4)
  CN kn_p{}, kn_pp{};  // stress coefficients of rows r-1, r-2
  CV kv_p{}, kv_pp{};  // velocity coefficients of rows r-2, r-3
  MV mv_p{}, mv_pp{};  // memory after step it, rows r-1, r-2
  MS ms_p{}, ms_pp{};  // memory after step it, rows r-2, r-3
  // receivers on the tile's own rows at all (one test per tile, not per row)
  const bool has_rec = nrec > 0 && S.rrow[ze] > S.rrow[zo];
  bool bad0 = false, bad1 = false;  // own cell above the blow-up bound at step it / it+1

  // Every load of iteration r (addresses clamped in bounds).  With
  // SDT_EL2X2_PREFETCH the loads of iteration r+1 are issued before
  // iteration r computes, so each wavefront keeps two rows in flight.
  struct Ld {
    T vx, vz, sxx, szz, sxz;
    CN kn;
    MV mv;
    CV kv;
    MS ms;
  };
  auto load_it = [&](int r) {
    Ld L;
    L.vx = vxi[cell(r + 1)];
    L.vz = vzi[cell(r + 1)];
    const uint32_t cA = cell(r);
    L.sxx = sxxi[cA];
    L.szz = szzi[cA];
    L.sxz = sxzi[cA];
    L.kn = load_cn(r);
    L.mv = load_mv(r);
    L.kv = load_cv(r - 1);
    L.ms = load_ms(r - 1);
    return L;
  };
  constexpr bool PF = SDT_EL2X2_PREFETCH && INT;
  Ld ahead{};
  if constexpr (PF) ahead = load_it(zo - 3);

  for (int r = zo - 3; r <= ze + 2; ++r) {
    const bool actA = r <= ze + 1, actB = r >= zo - 1 && r <= ze + 1;
    const bool actC = r >= zo + 1, actD = r >= zo + 3;
    Ld L;
    if constexpr (PF) {
      L = ahead;
      ahead = load_it(r + 1);
    } else {
      L = load_it(r);
    }
    const T ld_vx = L.vx, ld_vz = L.vz, ld_sxx = L.sxx, ld_szz = L.szz, ld_sxz = L.sxz;
    const CN kn = L.kn;
    MV mv = L.mv;
    const CV kv = L.kv;
    MS ms = L.ms;
    const bool nxt = r + 1 <= ze + 1 && valid(r + 1);
    const T vx0_n = nxt ? ld_vx : T(0), vz0_n = nxt ? ld_vz : T(0);

    // A: stresses of step it, row r
    const bool rvA = valid(r);
    T s1xx_c = rvA ? ld_sxx : T(0), s1zz_c = rvA ? ld_szz : T(0), s1xz_c = rvA ? ld_sxz : T(0);
    if (actA)
      stress(r, vx0_m, vz0_m, vx0_c, vz0_c, vx0_n, vz0_n, s1xx_c, s1zz_c, s1xz_c, kn, mv, ms0xx, ms0zz,
             ms0xz);
    // B: velocities of step it, row r-1
    T v1x_n = 0, v1z_n = 0;
    if (actB) {
      vel(r - 1, s1xx_p, s1zz_p, s1xz_p, s1xz_pp, s1zz_c, vx0_m, vz0_m, kv, ms, incx0, incz0, v1x_n, v1z_n);
      if (r - 1 >= zo && r - 1 < ze) {
        bad0 |= own && !((double)v1x_n * v1x_n + (double)v1z_n * v1z_n <= 1.0e50);
        if (has_rec) gather(r - 1, it, v1x_n, v1z_n);
      }
    }
    // C: stresses of step it+1, row r-2
    T s2xx_c = s1xx_pp, s2zz_c = s1zz_pp, s2xz_c = s1xz_pp;
    if (actC) {
      MV mc = mv_pp;
      stress(r - 2, v1x_m, v1z_m, v1x_c, v1z_c, v1x_n, v1z_n, s2xx_c, s2zz_c, s2xz_c, kn_pp, mc, ms1xx,
             ms1zz, ms1xz);
      if (own && r - 2 >= zo && r - 2 < ze) {
        const uint32_t c = cell(r - 2);
        sxxo[c] = s2xx_c;
        szzo[c] = s2zz_c;
        sxzo[c] = s2xz_c;
        store_mv(r - 2, mc);
      }
    }
    // D: velocities of step it+1, row r-3
    if (actD) {
      MS md = ms_pp;
      T v2x, v2z;
      vel(r - 3, s2xx_p, s2zz_p, s2xz_p, s2xz_pp, s2zz_c, v1x_m, v1z_m, kv_pp, md, incx1, incz1, v2x, v2z);
      if (r - 3 < ze) {
        bad1 |= own && !((double)v2x * v2x + (double)v2z * v2z <= 1.0e50);
        if (own) {
          const uint32_t c = cell(r - 3);
          vxo[c] = v2x;
          vzo[c] = v2z;
          store_ms(r - 3, md);
        }
        if (has_rec) gather(r - 3, it + 1, v2x, v2z);
      }
    }
    vx0_m = vx0_c; vz0_m = vz0_c; vx0_c = vx0_n; vz0_c = vz0_n;
    s1xx_pp = s1xx_p; s1zz_pp = s1zz_p; s1xz_pp = s1xz_p;
    s1xx_p = s1xx_c; s1zz_p = s1zz_c; s1xz_p = s1xz_c;
    v1x_m = v1x_c; v1z_m = v1z_c; v1x_c = v1x_n; v1z_c = v1z_n;
    s2xz_pp = s2xz_p; s2xx_p = s2xx_c; s2zz_p = s2zz_c; s2xz_p = s2xz_c;
    kn_pp = kn_p; kn_p = kn; kv_pp = kv_p; kv_p = kv;
    mv_pp = mv_p; mv_p = mv; ms_pp = ms_p; ms_p = ms;
  }
  if (bad0)
    check_blowup(HUGE_VAL, S.blowup, it);
  else if (bad1)
    check_blowup(HUGE_VAL, S.blowup, it + 1);
}

// Work: as k_el2_step, with EW2 owned columns per wavefront.  INT: every
// item of the work list is an interior block (el2x2_interior), so the kernel
// holds only the interior body and its smaller register budget allows more
// wavefronts per SIMD; otherwise each wavefront picks its body.
template <typename T, bool SPLIT, int WB, bool INT>
// (the split variant's C-PML body needs a larger budget: fewer wavefronts)
__global__ void __launch_bounds__(64 * WB, sizeof(T) == 4 ? (INT ? SDT_EL2X2_WAVES_INT : SPLIT ? SDT_EL2X2_WAVES_SPLIT : SDT_EL2X2_WAVES)
                                                          : (INT ? 3 : SPLIT ? 1 : 2))
k_el2_step2(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const int4* __restrict__ work, int nitems, const int* __restrict__ base, int rel,
            const int* __restrict__ it0, int eh0) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int it = __builtin_amdgcn_readfirstlane(*base + rel);
  const int q = __builtin_amdgcn_readfirstlane(((it - *it0) >> 1) & 1);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    const int4 wi = make_int4(__builtin_amdgcn_readfirstlane(wv.x),
                              __builtin_amdgcn_readfirstlane(wv.y) + wave * EW2,
                              __builtin_amdgcn_readfirstlane(wv.z), 0);
    if (wi.x < 0) continue;
    const DModel& M = models[shots[wi.x].model];
    if (wi.y >= M.nx) continue;
    const int eh = wv.w > 0 ? __builtin_amdgcn_readfirstlane(wv.w) : eh0;  // rows of this tile
    // the columns (lanes), resp. rows (zo-4 .. ze+1), the wave reads outside
    // the C-PML slabs and off the outer ring (columns 0, nx-1 and rows 0,
    // nz-1 and their neighbours)
    const int ze = min(wi.z + eh, M.nz);
    const bool xint = wi.y - 2 >= M.lx + 1 && wi.y + 61 <= M.nx - M.lx - 2;
    const bool zint = wi.z - 4 >= M.lz + 1 && ze + 1 <= M.nz - M.lz - 2;
    if (INT || (xint && zint))
      el2x2_tile<T, SPLIT, 0>(models, shots, wi, it, q, threadIdx.x & 63, eh);
    else if (xint)
      el2x2_tile<T, SPLIT, 1>(models, shots, wi, it, q, threadIdx.x & 63, eh);
    else if (zint)
      el2x2_tile<T, SPLIT, 3>(models, shots, wi, it, q, threadIdx.x & 63, eh);
    else
      el2x2_tile<T, SPLIT, 2>(models, shots, wi, it, q, threadIdx.x & 63, eh);
  }
}

// After a run whose last step wrote buffer B, copy B's fields and psi into A
// so the caller always finds the state in field[] / psi (A); B keeps an
// identical copy, which the next (odd) step reads.  With to_b the copy runs
// the other way (a run of two-step launches that ends in A before an odd
// step).
template <typename T>
__global__ void k_copy_back(const DShot* __restrict__ shots, int nfield, const long* __restrict__ nel,
                            const long* __restrict__ npsi, int to_b) {
  const int s = blockIdx.y;
  const DShot& S = shots[s];
  const long n = nel[s], np = npsi[s];
  const long tot = nfield * n + np;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
    if (q < nfield * n) {
      const int f = (int)(q / n);
      const long e = q - (long)f * n;
      if (to_b)
        mp<T>(S.g[f])[e] = cp<T>(S.f[f])[e];
      else
        mp<T>(S.f[f])[e] = cp<T>(S.g[f])[e];
    } else {
      const long e = q - nfield * n;
      if (to_b)
        mp<T>(S.psi2)[e] = cp<T>(S.psi)[e];
      else
        mp<T>(S.psi)[e] = cp<T>(S.psi2)[e];
    }
  }
}

// ===========================================================================
// ELASTIC2 adjoint: the transpose of one k_el2_step, for gradients of a
// trace misfit.  No counterpart in the reference (it has no adjoint).
//
// A step maps u_t = (v, sigma, psi) to u_{t+1} through the legacy sub-loops
// N (normal stresses), S (shear stress), V (velocities), source and
// Dirichlet edges, each linear in the state.  Its transpose runs them in
// reverse: with lambda_{t+1} = dJ/du_{t+1} (residual already injected at
// the receivers), adjoint step t produces lambda_t and adds
//   dJ/dcoef += lambda' * (d new state / d coef)
// where the partial products are recomputed from u_t and u_{t+1}, both
// read from the trail of recomputed forward states.  Three launches:
//   k_el2_adj_v  V^T : lambda of the new velocities -> partial lambda_v,
//                psi_s adjoint, and per-cell stress-derivative adjoints
//                (scratch planes 0..3);
//   k_el2_adj_s  gathers those into lambda_sigma, then S^T and N^T with
//                the shared velocity-derivative psi, writing per-cell
//                velocity-derivative adjoints (planes 4..11);
//   k_el2_adj_g  gathers those into lambda_v.
// Gradient slots are the kernel coefficient slots of the model
// (n11 n13 n15 n33 n35 s15 s35 s55 dt/rho_x dt/rho_z 1-gx 1-gz), FP64,
// accumulated atomically so shots of one model may share one array.
// ===========================================================================
__device__ __forceinline__ void gadd(double* g, double v) {
  if (v != 0.0) unsafeAtomicAdd(g, v);
}
// FP32 plans stage the gradient in FP32 within one sdt_adjoint_run call (at
// most a segment of steps) and flush it into the FP64 array at the end:
// half the atomic bytes, so a model's staging stays in L2.
__device__ __forceinline__ void gadd(float* g, double v) {
  const float f = (float)v;
  if (f != 0.0f) unsafeAtomicAdd(g, f);
}
__device__ __forceinline__ void gadd(__attribute__((address_space(1))) double* g, double v) {
  gadd((double*)g, v);
}
__device__ __forceinline__ void gadd(__attribute__((address_space(1))) float* g, double v) {
  gadd((float*)g, v);
}
// gradient products in FP32 (not FP64) before the FP32 staging add: a
// measured option, off by default
#ifndef SDT_ADJ_F32PROD
#define SDT_ADJ_F32PROD 0
#endif
#if SDT_ADJ_F32PROD
__device__ __forceinline__ void gadd(__attribute__((address_space(1))) float* g, float v) {
  if (v != 0.0f) unsafeAtomicAdd((float*)g, v);
}
#endif
template <typename T>
struct GAcc;
template <>
struct GAcc<double> {
  static __device__ __forceinline__ double* of(const DAdj& a) { return a.grad; }
};
template <>
struct GAcc<float> {
  static __device__ __forceinline__ float* of(const DAdj& a) { return a.g32; }
};

// grad += staging; staging = 0 (slots of the mask only)
__global__ void k_grad_flush(const DFlush* __restrict__ fl, unsigned gmask) {
  const DFlush& F = fl[blockIdx.y];
  for (int q = 0; q < 12; ++q) {
    if (!(gmask & (1u << q))) continue;
    double* g = F.grad + (long)q * F.n;
    float* h = F.g32 + (long)q * F.n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < F.n; e += (long)gridDim.x * blockDim.x) {
      g[e] += (double)h[e];
      h[e] = 0.0f;
    }
  }
}

// lambda_{t+1}(v) += dJ/dtrace(t) at the receivers (the trace of step t is
// read from u_{t+1}).
template <typename T>
__global__ void k_el2_adj_inject(const DModel* __restrict__ models, const DShot* __restrict__ shots,
                                 const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase,
                                 int k) {
  const int t = *tbase + k;
  const int s = blockIdx.y;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DAdj& A = adj[s];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= S.nrec || A.resid == nullptr) return;
  const DModel& M = models[S.model];
  const long n = (long)M.nx * M.nz;
  const auto q = cp<int>(S.rec) + 3 * r;
  const long c = (long)q[2] * M.nx + q[0];
  const auto res = cp<T>(A.resid) + (long)t * 2 * S.nrec + r;
  T* L = static_cast<T*>(A.lam[(t + 1) & 1]);
  atomicAdd(L + c, res[0]);
  atomicAdd(L + n + c, res[S.nrec]);
}

template <typename T, bool SPLIT>
__global__ void __launch_bounds__(NT)
k_el2_adj_v(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k,
            const int* __restrict__ tsel) {
  const int t = *tbase + k, t0 = *tbase;
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  const int lx = M.lx, lz = M.lz;
  const long SX = (long)nz * 2 * lx, SZ = (long)2 * lz * nx;
  const auto Lin = cp<T>(A.lam[(t + 1) & 1]);
  auto Lout = mp<T>(A.lam[t & 1]);
  const auto u1 = cp<T>(A.trail[*tsel]) + (long)(t + 1 - t0) * A.slot;
  const auto u0 = cp<T>(A.trail[*tsel]) + (long)(t - t0) * A.slot;
  auto W = mp<T>(A.work);
  double* G = A.grad;
  const Prof<T> P(&M.prof[0][0]);
  const bool inx = in_slab(i, nx, lx), inz = in_slab(j, nz, lz);
  const long ox = 5 * n + (inx ? (long)j * 2 * lx + slab_pos(i, nx, lx) : 0);
  const long oz = 5 * n + (SPLIT ? 6 : 4) * SX + (inz ? (long)slab_pos(j, nz, lz) * nx + i : 0);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz);
  // Dirichlet edges come last in the step: nothing flows back through them
  const bool edge = (i == 0 || i == nx - 1 || j == 0 || j == nz - 1);
  const T lvx = edge ? T(0) : Lin[c], lvz = edge ? T(0) : Lin[n + c];
  if (A.gsrc && i == S.isrc && j == S.ksrc) {  // d J / d source increment of step t
    A.gsrc[3 * (long)t + 0] = (double)lvx;
    A.gsrc[3 * (long)t + 2] = (double)lvz;
  }
  // vx(i,j) = vx * (1 - gx) + (dsxx_dx + dsxz_dz) * dt / rho, seismicfdtd.f95:317-337
  T ax1 = 0, ax2 = 0;
  if (i >= 1 && j >= 1) {
    const T w = cp<T>(M.coef[8])[c] * lvx;
    const T h1 = (u1[2 * n + c] - u1[2 * n + c - 1]) * idx;
    const T h2 = (u1[4 * n + c] - u1[4 * n + c - nx]) * idz;
    T b1 = h1, b2 = h2;
    ax1 = w;
    ax2 = w;
    if (inx) {
      const T lr = Lin[ox + 2 * SX] + w;
      ax1 = P[PK][i] * w + P[PA][i] * lr;
      Lout[ox + 2 * SX] = P[PB][i] * lr;
      b1 = P[PK][i] * h1 + u1[ox + 2 * SX];
    }
    if (inz) {
      const T lr = Lin[oz + 3 * SZ] + w;
      ax2 = P[12 + PK][j] * w + P[12 + PA][j] * lr;
      Lout[oz + 3 * SZ] = P[12 + PB][j] * lr;
      b2 = P[12 + PK][j] * h2 + u1[oz + 3 * SZ];
    }
    gadd(G + 8 * n + c, (double)lvx * (double)(b1 + b2));
    gadd(G + 10 * n + c, (double)lvx * (double)u0[c]);
    Lout[c] = (M.coef[10] ? cp<T>(M.coef[10])[c] : T(1)) * lvx;
  } else {
    if (inx) Lout[ox + 2 * SX] = Lin[ox + 2 * SX];
    if (inz) Lout[oz + 3 * SZ] = Lin[oz + 3 * SZ];
    Lout[c] = lvx;
  }
  // vz(i,j), seismicfdtd.f95:339-357
  T az1 = 0, az2 = 0;
  if (i <= nx - 2 && j <= nz - 2) {
    const T w = cp<T>(M.coef[9])[c] * lvz;
    const T k1 = (u1[4 * n + c + 1] - u1[4 * n + c]) * idx;
    const T k2 = (u1[3 * n + c + nx] - u1[3 * n + c]) * idz;
    T b1 = k1, b2 = k2;
    az1 = w;
    az2 = w;
    if (inx) {
      const T lr = Lin[ox + 3 * SX] + w;
      az1 = P[PKH][i] * w + P[PAH][i] * lr;
      Lout[ox + 3 * SX] = P[PBH][i] * lr;
      b1 = P[PKH][i] * k1 + u1[ox + 3 * SX];
    }
    if (inz) {
      const T lr = Lin[oz + 2 * SZ] + w;
      az2 = P[12 + PKH][j] * w + P[12 + PAH][j] * lr;
      Lout[oz + 2 * SZ] = P[12 + PBH][j] * lr;
      b2 = P[12 + PKH][j] * k2 + u1[oz + 2 * SZ];
    }
    gadd(G + 9 * n + c, (double)lvz * (double)(b1 + b2));
    gadd(G + 11 * n + c, (double)lvz * (double)u0[n + c]);
    Lout[n + c] = (M.coef[11] ? cp<T>(M.coef[11])[c] : T(1)) * lvz;
  } else {
    if (inx) Lout[ox + 3 * SX] = Lin[ox + 3 * SX];
    if (inz) Lout[oz + 2 * SZ] = Lin[oz + 2 * SZ];
    Lout[n + c] = lvz;
  }
  W[c] = ax1;          // adjoint of dsxx_dx at the vx node
  W[n + c] = ax2;      // adjoint of dsxz_dz at the vx node
  W[2 * n + c] = az1;  // adjoint of dsxz_dx at the vz node
  W[3 * n + c] = az2;  // adjoint of dszz_dz at the vz node
}

template <typename T, bool SPLIT>
__global__ void __launch_bounds__(NT)
k_el2_adj_s(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k,
            const int* __restrict__ tsel) {
  const int t = *tbase + k, t0 = *tbase;
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  const int lx = M.lx, lz = M.lz;
  const long SX = (long)nz * 2 * lx, SZ = (long)2 * lz * nx;
  const auto Lin = cp<T>(A.lam[(t + 1) & 1]);
  auto Lout = mp<T>(A.lam[t & 1]);
  const auto u1 = cp<T>(A.trail[*tsel]) + (long)(t + 1 - t0) * A.slot;
  const auto u0 = cp<T>(A.trail[*tsel]) + (long)(t - t0) * A.slot;
  auto W = mp<T>(A.work);
  double* G = A.grad;
  const Prof<T> P(&M.prof[0][0]);
  const bool inx = in_slab(i, nx, lx), inz = in_slab(j, nz, lz);
  const long ox = 5 * n + (inx ? (long)j * 2 * lx + slab_pos(i, nx, lx) : 0);
  const long oz = 5 * n + (SPLIT ? 6 : 4) * SX + (inz ? (long)slab_pos(j, nz, lz) * nx + i : 0);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dt = T(M.dt);
  // stresses feed the velocity update through their differences
  const T lsxx = Lin[2 * n + c] + idx * (W[c] - (i + 1 < nx ? W[c + 1] : T(0)));
  const T lszz = Lin[3 * n + c] + idz * ((j >= 1 ? W[3 * n + c - nx] : T(0)) - W[3 * n + c]);
  const T lsxz = Lin[4 * n + c] + idz * (W[n + c] - (j + 1 < nz ? W[n + c + nx] : T(0))) +
                 idx * ((i >= 1 ? W[2 * n + c - 1] : T(0)) - W[2 * n + c]);
  Lout[2 * n + c] = lsxx;
  Lout[3 * n + c] = lszz;
  Lout[4 * n + c] = lsxz;
  // adjoints of the velocity-derivative psi: shared by both stress loops
  // (A1), or the shear loop's own ones ls* in the split variant
  T lm1 = inx ? Lin[ox] : T(0), lm3 = inx ? Lin[ox + SX] : T(0);
  T lm2 = inz ? Lin[oz] : T(0), lm4 = inz ? Lin[oz + SZ] : T(0);
  T ls1 = 0, ls3 = 0, ls2 = 0, ls4 = 0;
  if (SPLIT) {
    ls1 = inx ? Lin[ox + 4 * SX] : T(0);
    ls3 = inx ? Lin[ox + 5 * SX] : T(0);
    ls2 = inz ? Lin[oz + 4 * SZ] : T(0);
    ls4 = inz ? Lin[oz + 5 * SZ] : T(0);
  }
  T& q1 = SPLIT ? ls1 : lm1;
  T& q2 = SPLIT ? ls2 : lm2;
  T& q3 = SPLIT ? ls3 : lm3;
  T& q4 = SPLIT ? ls4 : lm4;
  const long xq = SPLIT ? 4 * SX : 0, zq = SPLIT ? 4 * SZ : 0;  // shear-loop psi
  const T ikx = P[PK][i], ax = P[PA][i], bx = P[PB][i];
  const T ikxh = P[PKH][i], axh = P[PAH][i], bxh = P[PBH][i];
  const T ikz = P[12 + PK][j], az = P[12 + PA][j], bz = P[12 + PB][j];
  const T ikzh = P[12 + PKH][j], azh = P[12 + PAH][j], bzh = P[12 + PBH][j];
  // S^T: sigma_xz, seismicfdtd.f95:283-312 (psi after it is u_{t+1}'s)
  T lf1 = 0, lf2 = 0, lf3 = 0, lf4 = 0;
  if (j <= nz - 2 && i >= 1) {
    const T mu = dt * lsxz;
    const T f1 = (u0[c] - u0[c - 1]) * idx, f2 = (u0[n + c + nx] - u0[n + c]) * idz;
    const T f3 = (u0[n + c] - u0[n + c - 1]) * idx, f4 = (u0[c + nx] - u0[c]) * idz;
    const bool mono = M.coef[5] != nullptr;
    const T s15 = mono ? cp<T>(M.coef[5])[c] : T(0), s35 = mono ? cp<T>(M.coef[6])[c] : T(0);
    const T s55 = cp<T>(M.coef[7])[c];
    const T lg1 = s15 * mu, lg2 = s35 * mu, lg3 = s55 * mu;
    T g1 = f1, g2 = f2, g3 = f3, g4 = f4;
    lf1 = lg1; lf2 = lg2; lf3 = lg3; lf4 = lg3;
    if (inx) {
      g1 = ikxh * f1 + u1[ox + xq];
      T l = q1 + lg1;
      lf1 = ikxh * lg1 + axh * l;
      q1 = bxh * l;
      g3 = ikx * f3 + u1[ox + xq + SX];
      l = q3 + lg3;
      lf3 = ikx * lg3 + ax * l;
      q3 = bx * l;
    }
    if (inz) {
      g2 = ikz * f2 + u1[oz + zq];
      T l = q2 + lg2;
      lf2 = ikz * lg2 + az * l;
      q2 = bz * l;
      g4 = ikzh * f4 + u1[oz + zq + SZ];
      l = q4 + lg3;
      lf4 = ikzh * lg3 + azh * l;
      q4 = bzh * l;
    }
    const double m = (double)mu;
    gadd(G + 5 * n + c, m * (double)g1);
    gadd(G + 6 * n + c, m * (double)g2);
    gadd(G + 7 * n + c, m * ((double)g3 + (double)g4));
  }
  // N^T: sigma_xx, sigma_zz, seismicfdtd.f95:247-281 (psi before it is u_t's)
  T le1 = 0, le2 = 0, le3 = 0, le4 = 0;
  if (j >= 1 && i <= nx - 2) {
    const T mux = dt * lsxx, muz = dt * lszz;
    const T e1 = (u0[c + 1] - u0[c]) * idx, e2 = (u0[n + c] - u0[n + c - nx]) * idz;
    const T e3 = (u0[n + c + 1] - u0[n + c]) * idx, e4 = (u0[c] - u0[c - nx]) * idz;
    const bool mono = M.coef[2] != nullptr;
    const T n11 = cp<T>(M.coef[0])[c], n13 = cp<T>(M.coef[1])[c], n33 = cp<T>(M.coef[3])[c];
    const T n15 = mono ? cp<T>(M.coef[2])[c] : T(0), n35 = mono ? cp<T>(M.coef[4])[c] : T(0);
    const T ld1 = n11 * mux + n13 * muz, ld2 = n13 * mux + n33 * muz, lsh = n15 * mux + n35 * muz;
    T d1 = e1, d2 = e2, d3 = e3, d4 = e4;
    le1 = ld1; le2 = ld2; le3 = lsh; le4 = lsh;
    if (inx) {
      const T brow = SPLIT ? bxh : cp<T>(M.aux)[j];  // b_x_half(j), :255 (A2)
      d1 = ikxh * e1 + (brow * u0[ox] + axh * e1);
      T l = lm1 + ld1;
      le1 = ikxh * ld1 + axh * l;
      lm1 = brow * l;
      d3 = ikx * e3 + (bx * u0[ox + SX] + ax * e3);
      l = lm3 + lsh;
      le3 = ikx * lsh + ax * l;
      lm3 = bx * l;
    }
    if (inz) {
      d2 = ikz * e2 + (bz * u0[oz] + az * e2);
      T l = lm2 + ld2;
      le2 = ikz * ld2 + az * l;
      lm2 = bz * l;
      d4 = ikzh * e4 + (bzh * u0[oz + SZ] + azh * e4);
      l = lm4 + lsh;
      le4 = ikzh * lsh + azh * l;
      lm4 = bzh * l;
    }
    const double mx = (double)mux, mz = (double)muz, sh = (double)d3 + (double)d4;
    gadd(G + c, mx * (double)d1);
    gadd(G + n + c, mx * (double)d2 + mz * (double)d1);
    gadd(G + 2 * n + c, mx * sh);
    gadd(G + 3 * n + c, mz * (double)d2);
    gadd(G + 4 * n + c, mz * sh);
  }
  if (inx) {
    Lout[ox] = lm1;
    Lout[ox + SX] = lm3;
    if (SPLIT) {
      Lout[ox + 4 * SX] = ls1;
      Lout[ox + 5 * SX] = ls3;
    }
  }
  if (inz) {
    Lout[oz] = lm2;
    Lout[oz + SZ] = lm4;
    if (SPLIT) {
      Lout[oz + 4 * SZ] = ls2;
      Lout[oz + 5 * SZ] = ls4;
    }
  }
  W[4 * n + c] = le1;   // dvx_dx at the normal-stress node
  W[5 * n + c] = le2;   // dvz_dz
  W[6 * n + c] = le3;   // dvz_dx
  W[7 * n + c] = le4;   // dvx_dz
  W[8 * n + c] = lf1;   // dvx_dx at the shear node
  W[9 * n + c] = lf2;   // dvz_dz
  W[10 * n + c] = lf3;  // dvz_dx
  W[11 * n + c] = lf4;  // dvx_dz
}

template <typename T>
__global__ void __launch_bounds__(NT)
k_el2_adj_g(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k) {
  const int t = *tbase + k;
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  auto Lout = mp<T>(A.lam[t & 1]);
  const auto W = cp<T>(A.work);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz);
  const bool w_ = i >= 1, e_ = i + 1 < nx, n_ = j >= 1, s_ = j + 1 < nz;
  auto at = [&](int q, long o, bool ok) { return ok ? W[q * n + c + o] : T(0); };
  const T gx = idx * (at(4, -1, w_) - W[4 * n + c]) + idz * (W[7 * n + c] - at(7, nx, s_)) +
               idx * (W[8 * n + c] - at(8, 1, e_)) + idz * (at(11, -nx, n_) - W[11 * n + c]);
  const T gz = idz * (W[5 * n + c] - at(5, nx, s_)) + idx * (at(6, -1, w_) - W[6 * n + c]) +
               idz * (at(9, -nx, n_) - W[9 * n + c]) + idx * (W[10 * n + c] - at(10, 1, e_));
  Lout[c] = Lout[c] + gx;
  Lout[n + c] = Lout[n + c] + gz;
}

// ---------------------------------------------------------------------------
// Fused adjoint step: the same transpose as the three kernels above, in one
// pass.  A wavefront owns AW columns (lanes 2..AW+1; two halo lanes on each
// side) of AH rows of one shot and marches down them.  At march row r it
// does V^T of row r, the stress gathers with S^T / N^T of row r-1 and the
// velocity gather of row r-2, keeping the few rows each phase needs from
// the previous ones in registers and taking x neighbours from the
// neighbouring lanes.  The halo lanes and the two halo rows above and below
// the tile are recomputed, so tiles are independent (lambda ping-pongs).
// Only owned cells write lambda_t and add to the gradient.
// ---------------------------------------------------------------------------
#ifndef SDT_ADJ_WAVES_PER_SIMD
#define SDT_ADJ_WAVES_PER_SIMD 3
#endif
#ifndef SDT_ADJ_AW
#define SDT_ADJ_AW 60
#endif
constexpr int AW = SDT_ADJ_AW;   // owned columns per wavefront
static_assert(AW >= 1 && AW <= 60, "adjoint wavefront: two halo lanes on each side");
constexpr int AWB = 4;   // wavefronts per block, side by side in x
#ifndef SDT_ADJ_ROWS
#define SDT_ADJ_ROWS SDT_EL2_ROWS
#endif
constexpr int AH = SDT_ADJ_ROWS;

// One march row of the fused adjoint, shared by the one-step kernel and by
// both stages of the two-step kernel.  Inputs of march row r: lambda_{t+1}
// of the velocities of row r, of the stresses of row q = r - 1, the memory
// adjoints of row r (the stress-derivative memories the velocity update
// uses) and of row q (the velocity-derivative memories, with the shear
// loop's own four in the split variant).  Outputs: lambda_t of the
// velocities of row r - 2, of the stresses of row q, and the memory adjoints
// of rows r and q.
template <typename T>
struct AdjRow {
  T lvx = 0, lvz = 0, lsxx = 0, lszz = 0, lsxz = 0;
  T x2 = 0, x3 = 0, z2 = 0, z3 = 0;                                  // row r
  T x0 = 0, x1 = 0, z0 = 0, z1 = 0, sx0 = 0, sx1 = 0, sz0 = 0, sz1 = 0;  // row q
};

template <typename T, bool SPLIT, bool PML>
struct AdjStage {
  using GT = typename std::conditional<sizeof(T) == 4 && SDT_ADJ_F32PROD, float, double>::type;
  using GP = typename std::remove_pointer<decltype(GAcc<T>::of(std::declval<const DAdj&>()))>::type;
  const DShot& S;
  const DAdj& A;
  const int nx, nz, i, ic, lx, lz;
  const bool incol, sl, inx;
  const uint32_t n, SX, SZ, ZB, xs;
  const A32<const T> u1, u0;
  const A32<GP> G;
  T ikx, ax, bx, ikxh, axh, bxh, idx, idz, dt;
  gcptr<T> ikz, az, bz, ikzh, azh, bzh, brow;
  const A32<const T> N11, N13, N15, N33, N35, S15, S35, S55, DTX, DTZ, OMX, OMZ;
  const bool mono, atten;
  // windows (rows relative to the march row r)
  T A1 = 0, B1 = 0, C1 = 0, D1 = 0, D2 = 0;      // A, B, C, D of row r-1; D of r-2
  T pvx1 = 0, pvz1 = 0, pvx2 = 0, pvz2 = 0;      // partial lambda_v of rows r-1, r-2
  T vx_m = 0, vz_m = 0, vx_c = 0, vz_c = 0;      // u_t velocities of rows r-2, r-1
  T sxz_p = 0, szz_c = 0;                        // u_{t+1} sxz of row r-1, szz of row r
  T E1p = 0, E2p = 0, E3p = 0, E4p = 0, F1p = 0, F2p = 0, F3p = 0, F4p = 0;  // row r-2
  T F2pp = 0, F4pp = 0;                                                      // row r-3
  // per march row: rows, validity and memory offsets of rows r and q
  int rs = 0, qs = 0;
  bool vr = false, vq = false, inzr = false, inzq = false;
  uint32_t cr = 0, cq = 0, cn = 0, oxr = 0, oxq = 0, ozr = 0, ozq = 0;

  // u1 = the trail slot of u_{t+1}, u0 that of u_t; r0 = the first march row
  __device__ __forceinline__ AdjStage(const DModel& M, const DShot& S_, const DAdj& A_, int xo, int lane,
                                      const T* u1p, const T* u0p, int r0)
      : S(S_), A(A_), nx(M.nx), nz(M.nz), i(xo - 2 + lane),
        ic(i >= 0 && i < M.nx ? i : (i < 0 ? 0 : M.nx - 1)), lx(M.lx), lz(M.lz),
        incol(i >= 0 && i < M.nx), sl(incol && lane >= 1 && lane <= 2 + AW),
        inx(PML && incol && in_slab(i, M.nx, M.lx)), n((uint32_t)M.nx * M.nz),
        SX((uint32_t)M.nz * 2 * M.lx), SZ((uint32_t)2 * M.lz * M.nx),
        ZB(5 * n + (SPLIT ? 6 : 4) * SX), xs(5 * n + (inx ? slab_pos(i, M.nx, M.lx) : 0)),
        u1(u1p), u0(u0p), G(GAcc<T>::of(A_)), N11(M.coef[0]), N13(M.coef[1]), N15(M.coef[2]),
        N33(M.coef[3]), N35(M.coef[4]), S15(M.coef[5]), S35(M.coef[6]), S55(M.coef[7]),
        DTX(M.coef[8]), DTZ(M.coef[9]), OMX(M.coef[10]), OMZ(M.coef[11]),
        mono(M.coef[2] != nullptr), atten(M.coef[10] != nullptr) {
    const Prof<T> P(&M.prof[0][0]);
    ikx = P[PK][ic]; ax = P[PA][ic]; bx = P[PB][ic];
    ikxh = P[PKH][ic]; axh = P[PAH][ic]; bxh = P[PBH][ic];
    ikz = P[12 + PK]; az = P[12 + PA]; bz = P[12 + PB];
    ikzh = P[12 + PKH]; azh = P[12 + PAH]; bzh = P[12 + PBH];
    brow = cp<T>(M.aux);
    idx = T(1.0 / M.dx); idz = T(1.0 / M.dz); dt = T(M.dt);
    sxz_p = valid(r0 - 1) ? u1[4 * n + rowi(r0 - 1)] : T(0);
    szz_c = valid(r0) ? u1[3 * n + rowi(r0)] : T(0);
  }
  __device__ __forceinline__ uint32_t rowi(int r) const {
    return (uint32_t)(r < 0 ? 0 : (r >= nz ? nz - 1 : r)) * nx + ic;
  }
  __device__ __forceinline__ bool valid(int r) const { return incol && r >= 0 && r < nz; }
  __device__ __forceinline__ int clampr(int r) const { return r < 0 ? 0 : (r >= nz ? nz - 1 : r); }

  // rows, validity and memory offsets of march row r
  __device__ __forceinline__ void prep(int r) {
    const int q = r - 1;
    cr = rowi(r); cq = rowi(q); cn = rowi(r + 1);
    vr = valid(r); vq = valid(q);
    rs = clampr(r); qs = clampr(q);
    inzr = PML && incol && in_slab(rs, nz, lz);
    inzq = PML && incol && in_slab(qs, nz, lz);
    oxr = xs + (inx ? (uint32_t)rs * 2 * lx : 0);
    oxq = xs + (inx ? (uint32_t)qs * 2 * lx : 0);
    ozr = ZB + (inzr ? (uint32_t)slab_pos(rs, nz, lz) * nx + ic : 0);
    ozq = ZB + (inzq ? (uint32_t)slab_pos(qs, nz, lz) * nx + ic : 0);
  }

  // the inputs of the prepared row from a lambda state (slab loads only
  // where the cell has that memory: a zero-width slab has no storage)
  __device__ __forceinline__ void load(const A32<const T>& L, AdjRow<T>& in) const {
    in.lvx = L[cr]; in.lvz = L[n + cr];
    in.lsxx = L[2 * n + cq]; in.lszz = L[3 * n + cq]; in.lsxz = L[4 * n + cq];
    in.x2 = inx ? L[oxr + 2 * SX] : T(0); in.x3 = inx ? L[oxr + 3 * SX] : T(0);
    in.z2 = inzr ? L[ozr + 2 * SZ] : T(0); in.z3 = inzr ? L[ozr + 3 * SZ] : T(0);
    in.x0 = inx ? L[oxq] : T(0); in.x1 = inx ? L[oxq + SX] : T(0);
    in.z0 = inzq ? L[ozq] : T(0); in.z1 = inzq ? L[ozq + SZ] : T(0);
    if (SPLIT) {
      in.sx0 = inx ? L[oxq + 4 * SX] : T(0);
      in.sx1 = inx ? L[oxq + 5 * SX] : T(0);
      in.sz0 = inzq ? L[ozq + 4 * SZ] : T(0);
      in.sz1 = inzq ? L[ozq + 5 * SZ] : T(0);
    }
  }

  // One march row (prepared): V^T of row r, the stress gathers with S^T and
  // N^T of row q (from march row qmin on), the velocity gather of row r-2.
  // Gradient terms are added for the cells of rows [g0, g1) on the lanes
  // where `own`; the source gradient of step tg likewise.
  __device__ __forceinline__ void row(const int r, const AdjRow<T>& in, AdjRow<T>& o, const bool own,
                                      const int g0, const int g1, const int qmin, const int tg,
                                      const unsigned gmask) {
    const int q = r - 1;
    // ---- loads ----
    const T dtx = DTX[cr], dtz = DTZ[cr];
    const T omx = atten ? OMX[cr] : T(1), omz = atten ? OMZ[cr] : T(1);
    const T sxx1 = u1[2 * n + cr], sxz1 = u1[4 * n + cr], szz1n = u1[3 * n + cn];
    const T vx_n = vr ? u0[cr] : T(0), vz_n = vr ? u0[n + cr] : T(0);
    const T n11 = N11[cq], n13 = N13[cq], n33 = N33[cq], s55 = S55[cq];
    const T n15 = mono ? N15[cq] : T(0), n35 = mono ? N35[cq] : T(0);
    const T s15 = mono ? S15[cq] : T(0), s35 = mono ? S35[cq] : T(0);
    // psi of the velocity row r (stress-derivative memory) and of the stress
    // row q (velocity-derivative memory)
    const T Ux2 = inx ? u1[oxr + 2 * SX] : T(0), Ux3 = inx ? u1[oxr + 3 * SX] : T(0);
    const T Uz2 = inzr ? u1[ozr + 2 * SZ] : T(0), Uz3 = inzr ? u1[ozr + 3 * SZ] : T(0);
    const uint32_t xq2 = SPLIT ? 4 * SX : 0, zq2 = SPLIT ? 4 * SZ : 0;  // shear loop's arrays
    const T Ox0 = inx ? u0[oxq] : T(0), Ox1 = inx ? u0[oxq + SX] : T(0);
    const T Oz0 = inzq ? u0[ozq] : T(0), Oz1 = inzq ? u0[ozq + SZ] : T(0);
    const T Qx0 = inx ? u1[oxq + xq2] : T(0), Qx1 = inx ? u1[oxq + xq2 + SX] : T(0);
    const T Qz0 = inzq ? u1[ozq + zq2] : T(0), Qz1 = inzq ? u1[ozq + zq2 + SZ] : T(0);
    const T browq = SPLIT ? bxh : brow[qs];
    const T bzr = bz[rs], azr = az[rs], ikzr = ikz[rs], bzhr = bzh[rs], azhr = azh[rs], ikzhr = ikzh[rs];
    const T bzq = bz[qs], azq = az[qs], ikzq = ikz[qs], bzhq = bzh[qs], azhq = azh[qs], ikzhq = ikzh[qs];
    // neighbour values (wave-wide shuffles, outside any branch)
    const T sxx1_w = lane_prev(sxx1), sxz1_e = lane_next(sxz1);
    const T A1_e = lane_next(A1), C1_w = lane_prev(C1);
    const T vx_cw = lane_prev(vx_c), vz_cw = lane_prev(vz_c);
    const T vx_ce = lane_next(vx_c), vz_ce = lane_next(vz_c);
    const T E1p_w = lane_prev(E1p), E3p_w = lane_prev(E3p);
    const T F1p_e = lane_next(F1p), F3p_e = lane_next(F3p);

    // ---- V^T of row r: velocity updates, seismicfdtd.f95:317-375 ----
    const bool ownr = own && r >= g0 && r < g1;
    const bool edge = (i == 0 || i == nx - 1 || r == 0 || r == nz - 1);
    const T lvx = (vr && !edge) ? in.lvx : T(0), lvz = (vr && !edge) ? in.lvz : T(0);
    T An = 0, Bn = 0, Cn = 0, Dn = 0, pvx = lvx, pvz = lvz;
    if (ownr && i == S.isrc && r == S.ksrc && A.gsrc) {
      A.gsrc[3 * (long)tg + 0] = (double)lvx;
      A.gsrc[3 * (long)tg + 2] = (double)lvz;
    }
    o.x2 = in.x2; o.z3 = in.z3; o.x3 = in.x3; o.z2 = in.z2;
    if (vr && i >= 1 && r >= 1) {
      const T w = dtx * lvx;
      An = w;
      Bn = w;
      T b1 = (sxx1 - sxx1_w) * idx, b2 = (sxz1 - sxz_p) * idz;
      if (inx) {
        const T lr = in.x2 + w;
        An = ikx * w + ax * lr;
        o.x2 = bx * lr;
        b1 = ikx * b1 + Ux2;
      }
      if (inzr) {
        const T lr = in.z3 + w;
        Bn = ikzr * w + azr * lr;
        o.z3 = bzr * lr;
        b2 = ikzr * b2 + Uz3;
      }
      pvx = omx * lvx;
      if (ownr) {
        if (gmask & (1u << 8)) gadd(&G[8 * n + cr], (GT)lvx * (GT)(b1 + b2));
        if (gmask & (1u << 10)) gadd(&G[10 * n + cr], (GT)lvx * (GT)vx_n);
      }
    }
    if (vr && i <= nx - 2 && r <= nz - 2) {
      const T w = dtz * lvz;
      Cn = w;
      Dn = w;
      T b1 = (sxz1_e - sxz1) * idx, b2 = (szz1n - szz_c) * idz;
      if (inx) {
        const T lr = in.x3 + w;
        Cn = ikxh * w + axh * lr;
        o.x3 = bxh * lr;
        b1 = ikxh * b1 + Ux3;
      }
      if (inzr) {
        const T lr = in.z2 + w;
        Dn = ikzhr * w + azhr * lr;
        o.z2 = bzhr * lr;
        b2 = ikzhr * b2 + Uz2;
      }
      pvz = omz * lvz;
      if (ownr) {
        if (gmask & (1u << 9)) gadd(&G[9 * n + cr], (GT)lvz * (GT)(b1 + b2));
        if (gmask & (1u << 11)) gadd(&G[11 * n + cr], (GT)lvz * (GT)vz_n);
      }
    }

    // ---- stresses of row q: gathers, then S^T and N^T, :247-312 ----
    T Ee1 = 0, Ee2 = 0, Ee3 = 0, Ee4 = 0, Ff1 = 0, Ff2 = 0, Ff3 = 0, Ff4 = 0;
    o.lsxx = 0; o.lszz = 0; o.lsxz = 0;
    o.x0 = in.x0; o.x1 = in.x1; o.z0 = in.z0; o.z1 = in.z1;
    o.sx0 = in.sx0; o.sx1 = in.sx1; o.sz0 = in.sz0; o.sz1 = in.sz1;
    if (sl && vq && q >= qmin) {
      const bool ownq = own && q >= g0 && q < g1;
      const T lsxx = in.lsxx + idx * (A1 - A1_e);
      const T lszz = in.lszz + idz * (D2 - D1);
      const T lsxz = in.lsxz + idz * (B1 - Bn) + idx * (C1_w - C1);
      o.lsxx = lsxx; o.lszz = lszz; o.lsxz = lsxz;
      T lm1 = in.x0, lm3 = in.x1, lm2 = in.z0, lm4 = in.z1;     // normal loop's psi adjoints
      T ls1 = in.sx0, ls3 = in.sx1, ls2 = in.sz0, ls4 = in.sz1;  // shear loop's (split)
      T& k1 = SPLIT ? ls1 : lm1;
      T& k2 = SPLIT ? ls2 : lm2;
      T& k3 = SPLIT ? ls3 : lm3;
      T& k4 = SPLIT ? ls4 : lm4;
      if (q <= nz - 2 && i >= 1) {  // S^T
        const T mu = dt * lsxz;
        const T f1 = (vx_c - vx_cw) * idx, f2 = (vz_n - vz_c) * idz;
        const T f3 = (vz_c - vz_cw) * idx, f4 = (vx_n - vx_c) * idz;
        const T lg1 = s15 * mu, lg2 = s35 * mu, lg3 = s55 * mu;
        T g1 = f1, g2 = f2, g3 = f3, g4 = f4;
        Ff1 = lg1; Ff2 = lg2; Ff3 = lg3; Ff4 = lg3;
        if (inx) {
          g1 = ikxh * f1 + Qx0;
          T l = k1 + lg1;
          Ff1 = ikxh * lg1 + axh * l;
          k1 = bxh * l;
          g3 = ikx * f3 + Qx1;
          l = k3 + lg3;
          Ff3 = ikx * lg3 + ax * l;
          k3 = bx * l;
        }
        if (inzq) {
          g2 = ikzq * f2 + Qz0;
          T l = k2 + lg2;
          Ff2 = ikzq * lg2 + azq * l;
          k2 = bzq * l;
          g4 = ikzhq * f4 + Qz1;
          l = k4 + lg3;
          Ff4 = ikzhq * lg3 + azhq * l;
          k4 = bzhq * l;
        }
        if (ownq) {
          const GT m = (GT)mu;
          if (gmask & (1u << 5)) gadd(&G[5 * n + cq], m * (GT)g1);
          if (gmask & (1u << 6)) gadd(&G[6 * n + cq], m * (GT)g2);
          if (gmask & (1u << 7)) gadd(&G[7 * n + cq], m * ((GT)g3 + (GT)g4));
        }
      }
      if (q >= 1 && i <= nx - 2) {  // N^T
        const T mux = dt * lsxx, muz = dt * lszz;
        const T e1 = (vx_ce - vx_c) * idx, e2 = (vz_c - vz_m) * idz;
        const T e3 = (vz_ce - vz_c) * idx, e4 = (vx_c - vx_m) * idz;
        const T ld1 = n11 * mux + n13 * muz, ld2 = n13 * mux + n33 * muz, lsh = n15 * mux + n35 * muz;
        T d1 = e1, d2 = e2, d3 = e3, d4 = e4;
        Ee1 = ld1; Ee2 = ld2; Ee3 = lsh; Ee4 = lsh;
        if (inx) {
          d1 = ikxh * e1 + (browq * Ox0 + axh * e1);
          T l = lm1 + ld1;
          Ee1 = ikxh * ld1 + axh * l;
          lm1 = browq * l;
          d3 = ikx * e3 + (bx * Ox1 + ax * e3);
          l = lm3 + lsh;
          Ee3 = ikx * lsh + ax * l;
          lm3 = bx * l;
        }
        if (inzq) {
          d2 = ikzq * e2 + (bzq * Oz0 + azq * e2);
          T l = lm2 + ld2;
          Ee2 = ikzq * ld2 + azq * l;
          lm2 = bzq * l;
          d4 = ikzhq * e4 + (bzhq * Oz1 + azhq * e4);
          l = lm4 + lsh;
          Ee4 = ikzhq * lsh + azhq * l;
          lm4 = bzhq * l;
        }
        if (ownq) {
          const GT mx = (GT)mux, mz = (GT)muz, sh = (GT)d3 + (GT)d4;
          if (gmask & (1u << 0)) gadd(&G[cq], mx * (GT)d1);
          if (gmask & (1u << 1)) gadd(&G[n + cq], mx * (GT)d2 + mz * (GT)d1);
          if (gmask & (1u << 2)) gadd(&G[2 * n + cq], mx * sh);
          if (gmask & (1u << 3)) gadd(&G[3 * n + cq], mz * (GT)d2);
          if (gmask & (1u << 4)) gadd(&G[4 * n + cq], mz * sh);
        }
      }
      o.x0 = lm1; o.x1 = lm3; o.z0 = lm2; o.z1 = lm4;
      o.sx0 = ls1; o.sx1 = ls3; o.sz0 = ls2; o.sz1 = ls4;
    }

    // ---- velocities of row r-2: gather of the stress-phase adjoints ----
    o.lvx = pvx2 + idx * (E1p_w - E1p) + idz * (E4p - Ee4) + idx * (F1p - F1p_e) + idz * (F4pp - F4p);
    o.lvz = pvz2 + idz * (E2p - Ee2) + idx * (E3p_w - E3p) + idz * (F2pp - F2p) + idx * (F3p - F3p_e);

    // ---- shift the windows ----
    F2pp = F2p; F4pp = F4p;
    E1p = Ee1; E2p = Ee2; E3p = Ee3; E4p = Ee4;
    F1p = Ff1; F2p = Ff2; F3p = Ff3; F4p = Ff4;
    pvx2 = pvx1; pvz2 = pvz1; pvx1 = pvx; pvz1 = pvz;
    A1 = An; B1 = Bn; C1 = Cn; D2 = D1; D1 = Dn;
    vx_m = vx_c; vz_m = vz_c; vx_c = vx_n; vz_c = vz_n;
    sxz_p = sxz1; szz_c = szz1n;
  }

  // Store the outputs of the prepared march row r into a lambda state for
  // the cells of rows [g0, g1) on the lanes where `own`.
  __device__ __forceinline__ void store(const A32<T>& L, const int r, const AdjRow<T>& o, const bool own,
                                        const int g0, const int g1) const {
    const int q = r - 1, pr = r - 2;
    if (own && r >= g0 && r < g1) {
      if (inx) { L[oxr + 2 * SX] = o.x2; L[oxr + 3 * SX] = o.x3; }
      if (inzr) { L[ozr + 3 * SZ] = o.z3; L[ozr + 2 * SZ] = o.z2; }
    }
    if (own && q >= g0 && q < g1) {
      L[2 * n + cq] = o.lsxx;
      L[3 * n + cq] = o.lszz;
      L[4 * n + cq] = o.lsxz;
      if (inx) {
        L[oxq] = o.x0;
        L[oxq + SX] = o.x1;
        if (SPLIT) {
          L[oxq + 4 * SX] = o.sx0;
          L[oxq + 5 * SX] = o.sx1;
        }
      }
      if (inzq) {
        L[ozq] = o.z0;
        L[ozq + SZ] = o.z1;
        if (SPLIT) {
          L[ozq + 4 * SZ] = o.sz0;
          L[ozq + 5 * SZ] = o.sz1;
        }
      }
    }
    if (own && pr >= g0 && pr < g1) {
      const uint32_t cp_ = rowi(pr);
      L[cp_] = o.lvx;
      L[n + cp_] = o.lvz;
    }
  }
};

template <typename T, bool SPLIT, bool PML>
__device__ __forceinline__ void el2_adj_tile(const DModel* __restrict__ models,
                                             const DShot* __restrict__ shots,
                                             const DAdj* __restrict__ adj, const int s, const int xo,
                                             const int zo, const int t, const int t0,
                                             const unsigned gmask, const int lane, const int tb) {
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int zend = min(zo + AH, M.nz);
  const A32<const T> Lin(A.lam[(t + 1) & 1]);
  const A32<T> Lout(A.lam[t & 1]);
  const T* trail = static_cast<const T*>(A.trail[tb]);
  AdjStage<T, SPLIT, PML> st(M, S, A, xo, lane, trail + (long)(t + 1 - t0) * A.slot,
                             trail + (long)(t - t0) * A.slot, zo - 2);
  const bool own = st.incol && lane >= 2 && lane < 2 + AW;
  for (int r = zo - 2; r <= zend + 1; ++r) {
    st.prep(r);
    AdjRow<T> in, o;
    st.load(Lin, in);
    st.row(r, in, o, own, zo, zend, zo - 1, t, gmask);
    st.store(Lout, r, o, own, zo, zend);
  }
}

template <typename T, bool SPLIT>
__global__ void __launch_bounds__(64 * AWB, sizeof(T) == 4 ? SDT_ADJ_WAVES_PER_SIMD : 2)
k_el2_adj(const DModel* __restrict__ models, const DShot* __restrict__ shots,
          const DAdj* __restrict__ adj, const int4* __restrict__ work, int nitems,
          const int* __restrict__ tbase, int k, unsigned gmask, const int* __restrict__ tsel) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t0 = __builtin_amdgcn_readfirstlane(*tbase);
  const int tb = __builtin_amdgcn_readfirstlane(*tsel);
  const int t = t0 + k;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    const int s = __builtin_amdgcn_readfirstlane(wv.x);
    const int xo = __builtin_amdgcn_readfirstlane(wv.y) + wave * AW;
    const int zo = __builtin_amdgcn_readfirstlane(wv.z);
    if (s < 0) continue;
    const DModel& M = models[shots[s].model];
    if (xo >= M.nx) continue;
    // a wave whose columns and rows (halo included) avoid every C-PML slab
    // runs the body without memory variables
    const bool interior = xo - 2 >= M.lx && xo + 61 < M.nx - M.lx && zo - 3 >= M.lz &&
                          zo + AH + 2 < M.nz - M.lz;
    if (interior)
      el2_adj_tile<T, SPLIT, false>(models, shots, adj, s, xo, zo, t, t0, gmask, threadIdx.x & 63, tb);
    else
      el2_adj_tile<T, SPLIT, true>(models, shots, adj, s, xo, zo, t, t0, gmask, threadIdx.x & 63, tb);
  }
}

// ===========================================================================
// EM2: electromag2, legacy/fortran/electromagfdtd.f95:10-326, one pass per
// step.  psi slabs: x [nz][2lx]: dEz_dx dHy_dx ; z [2lz][nx]: dEx_dz dHy_dz.
// fields: Ex (nx-1 x nz), Ez (nx x nz-1), Hy (nx-1 x nz-1) stored nx*nz.
// A wavefront owns columns xo..xo+EW-1 (lanes 1..EW) of rows zo..zo+eh-1 and
// walks down z: at row r it updates Hy of row r from the old Ex of rows r and
// r+1 and the old Ez of columns i and i+1 (:249-266), then Ex and Ez of row r
// from the new Hy of rows r-1 and r and of columns i-1 and i (:273-295), then
// the source (:300-301) and the Dirichlet edges (:304-311).  The new Hy of
// row zo-1 is recomputed, so tiles are independent; the state ping-pongs
// between buffers A and B like the 2-D elastic one.  The Hy zeroing that ends
// each legacy step (:312-315) is applied by the next update, which reads a
// boundary Hy as 0, and by k_em2_edges at the end of a run.
// Every fused multiply-add is written out (contraction off).
// ===========================================================================
template <typename T, bool TRAIL, bool PML>
__device__ __forceinline__ void em2_tile(const DModel* __restrict__ models,
                                         const DShot* __restrict__ shots, const int4 wi, const int it,
                                         const int lane, const DAdj* __restrict__ adj, const int slot,
                                         const int tb, const int eh, const int par) {
#pragma clang fp contract(off)
  const DShot& S = shots[wi.x];
  const DModel& M = models[S.model];
  const int nx = M.nx, nz = M.nz;
  const int xo = wi.y, zo = wi.z;
  const int i = xo - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);  // clamped column: every load stays in bounds
  const bool own = incol && lane >= 1 && lane <= EW;
  const int zend = min(zo + eh, nz);
  const bool has_rec = !TRAIL && S.nrec > 0 && S.rrow[zend] > S.rrow[zo];  // receivers on own rows
  const long nn = (long)nx * nz;
  const T* tin = nullptr;
  T* tout = nullptr;
  if (TRAIL) {
    const DAdj& A = adj[wi.x];
    tin = static_cast<const T*>(A.trail[tb]) + (long)slot * A.slot;
    tout = static_cast<T*>(A.trail[tb]) + (long)(slot + 1) * A.slot;
  }
  const A32<const T> exi(TRAIL ? tin : (par ? S.g[0] : S.f[0]));
  const A32<const T> ezi(TRAIL ? tin + nn : (par ? S.g[1] : S.f[1]));
  const A32<const T> hyi(TRAIL ? tin + 2 * nn : (par ? S.g[2] : S.f[2]));
  const A32<T> exo(TRAIL ? tout : (par ? S.f[0] : S.g[0]));
  const A32<T> ezo(TRAIL ? tout + nn : (par ? S.f[1] : S.g[1]));
  const A32<T> hyo(TRAIL ? tout + 2 * nn : (par ? S.f[2] : S.g[2]));
  const A32<const T> pin(TRAIL ? tin + 3 * nn : (par ? S.psi2 : S.psi));
  const A32<T> pout(TRAIL ? tout + 3 * nn : (par ? S.psi : S.psi2));
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dbh = T(M.scal[0]);
  const int lx = M.lx, lz = M.lz;
  const uint32_t SX = (uint32_t)nz * 2 * lx, SZ = (uint32_t)2 * lz * nx;
  const bool inx = PML && incol && in_slab(i, nx, lx);
  const int xs = inx ? slab_pos(i, nx, lx) : 0;
  const Prof<T> P(&M.prof[0][0]);
  const T ikx = P[PK][ic], ax = P[PA][ic], bx = P[PB][ic];
  const T ikxh = P[PKH][ic], axh = P[PAH][ic], bxh = P[PBH][ic];
  const auto ikz = P[12 + PK], az = P[12 + PA], bz = P[12 + PB];
  const auto ikzh = P[12 + PKH], azh = P[12 + PAH], bzh = P[12 + PBH];
  const A32<const T> CAX(M.coef[0]), CBX(M.coef[1]), CAZ(M.coef[2]), CBZ(M.coef[3]);
  const bool hcol = incol && i <= nx - 2;                  // Hy, Ex columns
  const bool ezcol = incol && i >= 1 && i <= nx - 2;       // Ez columns
  const bool is_src_col = (i == S.isrc);
  const int ksrc = S.ksrc;
  const auto srcx = cp<T>(S.src[0]), srcz = cp<T>(S.src[2]);
  const T incx = (srcx != nullptr) ? srcx[it] : T(0);
  const T incz = (srcz != nullptr) ? srcz[it] : T(0);

  auto row = [&](int r) { return (uint32_t)(r < 0 ? 0 : (r >= nz ? nz - 1 : r)) * nx + ic; };
  auto valid = [&](int r) { return incol && r >= 0 && r < nz; };
  T ex_c = valid(zo - 1) ? exi[row(zo - 1)] : T(0);  // old Ex of row r
  T hy_p = 0;                                         // new Hy of row r-1
  for (int r = zo - 1; r < zend; ++r) {
    const uint32_t c = row(r);
    const bool rv = valid(r);
    const T ld_exn = exi[row(r + 1)], ld_ez = ezi[c], ld_hy = hyi[c];
    const T cax = CAX[c], cbx = CBX[c], caz = CAZ[c], cbz = CBZ[c];
    const int rs = r < 0 ? 0 : (r >= nz ? nz - 1 : r);
    const bool inz = PML && in_slab(rs, nz, lz), inzc = incol && inz;
    const uint32_t ox = (uint32_t)rs * 2 * lx + xs;
    const uint32_t oz = 2 * SX + (inz ? (uint32_t)slab_pos(rs, nz, lz) * nx + ic : 0);
    T m_ezx = inx ? pin[ox] : T(0);        // dEz_dx
    T m_hyx = inx ? pin[SX + ox] : T(0);   // dHy_dx
    T m_exz = inzc ? pin[oz] : T(0);       // dEx_dz
    T m_hyz = inzc ? pin[SZ + oz] : T(0);  // dHy_dz
    const T ex_n = valid(r + 1) ? ld_exn : T(0);
    const T ez_c = rv ? ld_ez : T(0);
    const T hy_o = rv ? ld_hy : T(0);

    // ---- Hy of row r (:249-266) ----
    const T ez_e = lane_next(ez_c);  // Ez(i+1, r)
    T hy = hy_o;
    if (rv && hcol && r <= nz - 2) {
      T a = (ex_n - ex_c) * idz;
      if (inz) {
        m_exz = fma(bz[rs], m_exz, az[rs] * a);
        a = fma(a, ikz[rs], m_exz);
      }
      T b = (ez_e - ez_c) * idx;
      if (inx) {
        m_ezx = fma(bx, m_ezx, ax * b);
        b = fma(b, ikx, m_ezx);
      }
      const bool edge = (i == 0 || i == nx - 2 || r == 0 || r == nz - 2);  // :312-315 of the last step
      hy = fma(dbh, b + a, edge ? T(0) : hy_o);
    }

    // ---- Ex, Ez of row r (:273-295), source, Dirichlet ----
    const T hy_w = lane_prev(hy);  // new Hy(i-1, r)
    T gex = 0, gez = 0;  // new E of the owned cells (receiver samples)
    if (own && rv && r >= zo) {
      T ex = ex_c, ez = ez_c;
      if (hcol && r >= 1 && r <= nz - 2) {
        T d = (hy - hy_p) * idz;
        if (inz) {
          m_hyz = fma(bzh[rs], m_hyz, azh[rs] * d);
          d = fma(d, ikzh[rs], m_hyz);
        }
        ex = fma(cax, ex, cbx * d);
      }
      if (ezcol && r <= nz - 2) {
        T d = (hy - hy_w) * idx;
        if (inx) {
          m_hyx = fma(bxh, m_hyx, axh * d);
          d = fma(d, ikxh, m_hyx);
        }
        ez = fma(caz, ez, cbz * d);
      }
      if (is_src_col && r == ksrc) { ex += incx; ez += incz; }
      if (i == 0 || i == nx - 2 || r == 0 || r == nz - 1) ex = 0;
      if (i == 0 || i == nx - 1 || r == 0 || r == nz - 2) ez = 0;
      if (i == nx - 1) ex = 0;  // outside the (nx-1) x nz Ex array
      if (r == nz - 1) ez = 0;  // outside the nx x (nz-1) Ez array
      check_blowup((double)ex * ex + (double)ez * ez, S.blowup, it);
      exo[c] = ex;
      ezo[c] = ez;
      hyo[c] = hy;
      gex = ex;
      gez = ez;
      if (inx) { pout[ox] = m_ezx; pout[SX + ox] = m_hyx; }
      if (inz) { pout[oz] = m_exz; pout[SZ + oz] = m_hyz; }
    }
    if (has_rec && r >= zo) gather_row<T>(S, r, it, lane, xo - 1, 1, EW, gex, gez);
    hy_p = hy;
    ex_c = ex_n;
  }
}

// Same work list and launch shape as k_el2_step, and the same buffer
// choice (prun: the single step that ends a run of two-step launches).
template <typename T, bool TRAIL, int WB>
__global__ void __launch_bounds__(64 * WB)
k_em2_step(const DModel* __restrict__ models, const DShot* __restrict__ shots,
           const int4* __restrict__ work, int nitems, const int* __restrict__ base, int rel,
           const DAdj* __restrict__ adj, const int* __restrict__ tsel, int eh,
           const int* __restrict__ prun) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int it = __builtin_amdgcn_readfirstlane(*base + rel);
  const int par = prun ? __builtin_amdgcn_readfirstlane(((it - *prun) >> 1) & 1) : (it & 1);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    const int4 wi = make_int4(__builtin_amdgcn_readfirstlane(wv.x),
                              __builtin_amdgcn_readfirstlane(wv.y) + wave * EW,
                              __builtin_amdgcn_readfirstlane(wv.z), 0);
    if (wi.x < 0) continue;
    const DModel& M = models[shots[wi.x].model];
    if (wi.y >= M.nx) continue;
    // columns (halo included) and rows zo-1 .. zo+eh-1 clear of every slab
    const bool interior = wi.y - 1 >= M.lx && wi.y + 62 < M.nx - M.lx && wi.z - 1 >= M.lz &&
                          wi.z + eh <= M.nz - M.lz;
    const int tb = TRAIL ? __builtin_amdgcn_readfirstlane(*tsel) : 0;
    if (interior)
      em2_tile<T, TRAIL, false>(models, shots, wi, it, threadIdx.x & 63, adj, rel, tb, eh, par);
    else
      em2_tile<T, TRAIL, true>(models, shots, wi, it, threadIdx.x & 63, adj, rel, tb, eh, par);
  }
}

// ===========================================================================
// EM2, two time steps per launch: the walk of k_em2_step with the second
// step one row behind the first.  At row r a wavefront computes
//   A  Hy of step it at row r        (E of it-1, rows r, r+1)
//   B  Ex, Ez of step it at row r     (Hy from A, rows r-1, r)
//   C  Hy of step it+1 at row r-1     (E from B, rows r-1, r)
//   D  Ex, Ez of step it+1 at row r-1 (Hy from C, rows r-2, r-1)
// so the state of step it lives only in registers.  One step makes a cell
// depend on columns i-1 .. i+1 and rows k-1 .. k+1 of the previous state,
// so lane l holds column xo - 2 + l, lanes 2 .. 2+EW2E-1 are exact after two
// steps, and rows zo-2 .. zo+eh read produce rows zo .. zo+eh-1.  The
// expressions are those of k_em2_step in the same order (identical bits),
// the Hy zeroing that ends step it is applied where step it+1 reads it, and
// the receiver samples of both steps are written here.  Buffers as in
// k_el2_step2.
// ===========================================================================
constexpr int EW2E = 60;

template <typename T, bool PML>
__device__ __forceinline__ void em2x2_tile(const DModel* __restrict__ models,
                                           const DShot* __restrict__ shots, const int4 wi, const int it,
                                           const int q, const int lane, const int eh) {
#pragma clang fp contract(off)
  // interior waves read only columns 1..nx-3 and rows 1..nz-3 outside the
  // C-PML slabs: every edge condition holds on every lane
  constexpr bool INT = !PML;
  const DShot& S = shots[wi.x];
  const DModel& M = models[S.model];
  const int nx = M.nx, nz = M.nz;
  const int xo = wi.y, zo = wi.z;
  const int i = xo - 2 + lane;
  const bool incol = INT || (i >= 0 && i < nx);
  const int ic = INT ? i : (incol ? i : (i < 0 ? 0 : nx - 1));  // clamped column
  const bool own = incol && lane >= 2 && lane < 2 + EW2E;
  const int ze = min(zo + eh, nz);
  const A32<const T> exi(q ? S.g[0] : S.f[0]);
  const A32<const T> ezi(q ? S.g[1] : S.f[1]);
  const A32<const T> hyi(q ? S.g[2] : S.f[2]);
  const A32<T> exo(q ? S.f[0] : S.g[0]);
  const A32<T> ezo(q ? S.f[1] : S.g[1]);
  const A32<T> hyo(q ? S.f[2] : S.g[2]);
  const A32<const T> pin(q ? S.psi2 : S.psi);
  const A32<T> pout(q ? S.psi : S.psi2);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dbh = T(M.scal[0]);
  const int lx = M.lx, lz = M.lz;
  const uint32_t SX = (uint32_t)nz * 2 * lx, SZ = (uint32_t)2 * lz * nx;
  const bool inx = PML && incol && in_slab(i, nx, lx);
  const int xs = inx ? slab_pos(i, nx, lx) : 0;
  const Prof<T> P(&M.prof[0][0]);
  const T ikx = P[PK][ic], ax = P[PA][ic], bx = P[PB][ic];
  const T ikxh = P[PKH][ic], axh = P[PAH][ic], bxh = P[PBH][ic];
  const auto ikz = kp<T>(M.prof[2][PK]), az = kp<T>(M.prof[2][PA]), bz = kp<T>(M.prof[2][PB]);
  const auto ikzh = kp<T>(M.prof[2][PKH]), azh = kp<T>(M.prof[2][PAH]), bzh = kp<T>(M.prof[2][PBH]);
  const A32<const T> CAX(M.coef[0]), CBX(M.coef[1]), CAZ(M.coef[2]), CBZ(M.coef[3]);
  const bool hcol = INT || (incol && i <= nx - 2);              // Hy, Ex columns
  const bool ezcol = INT || (incol && i >= 1 && i <= nx - 2);   // Ez columns
  const bool is_src_col = (i == S.isrc);
  const int ksrc = S.ksrc;
  const auto srcx = cp<T>(S.src[0]), srcz = cp<T>(S.src[2]);
  const T incx0 = srcx ? srcx[it] : T(0), incx1 = srcx ? srcx[it + 1] : T(0);
  const T incz0 = srcz ? srcz[it] : T(0), incz1 = srcz ? srcz[it + 1] : T(0);

  auto crow = [&](int r) { return r < 0 ? 0 : (r >= nz ? nz - 1 : r); };
  auto cell = [&](int r) { return (uint32_t)crow(r) * nx + ic; };
  auto valid = [&](int r) { return INT || (incol && r >= 0 && r < nz); };
  auto inzs = [&](int r) { return PML && in_slab(crow(r), nz, lz); };

  struct MX { T ezx, hyx, exz, hyz; };  // memory: dEz/dx, dHy/dx (x slab); dEx/dz, dHy/dz (z slab)
  struct CE { T cax, cbx, caz, cbz; };
  auto offx = [&](int r) { return inx ? (uint32_t)crow(r) * 2 * lx + xs : 0u; };
  auto offz = [&](int r) { return 2 * SX + (inzs(r) ? (uint32_t)slab_pos(crow(r), nz, lz) * nx + ic : 0u); };
  auto load_m = [&](int r) {
    MX m{};
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      const bool inzc = incol && inzs(r);
      m.ezx = inx ? pin[ox] : T(0);
      m.hyx = inx ? pin[SX + ox] : T(0);
      m.exz = inzc ? pin[oz] : T(0);
      m.hyz = inzc ? pin[SZ + oz] : T(0);
    }
    return m;
  };
  auto store_m = [&](int r, const MX& m) {  // own cells only
    if constexpr (PML) {
      const uint32_t ox = offx(r), oz = offz(r);
      if (inx) { pout[ox] = m.ezx; pout[SX + ox] = m.hyx; }
      if (inzs(r)) { pout[oz] = m.exz; pout[SZ + oz] = m.hyz; }
    }
  };
  // Hy of one step at row r (electromagfdtd.f95:249-266): the previous step's
  // Ex of rows r, r+1, Ez of row r (columns i, i+1) and Hy of row r
  auto hfield = [&](int r, T ex_c, T ex_n, T ez_c, T hy_o, MX& m) {
    const T ez_e = dpp_next(ez_c);  // Ez(i+1, r)
    const bool rv = valid(r);
    const int rs = crow(r);
    const bool inz = PML && in_slab(rs, nz, lz);
    T hy = hy_o;
    if (INT || (rv && hcol && r <= nz - 2)) {
      T a = (ex_n - ex_c) * idz;
      if (PML && inz) {
        m.exz = fma(bz[rs], m.exz, az[rs] * a);
        a = fma(a, ikz[rs], m.exz);
      }
      T b = (ez_e - ez_c) * idx;
      if (PML && inx) {
        m.ezx = fma(bx, m.ezx, ax * b);
        b = fma(b, ikx, m.ezx);
      }
      const bool edge = !INT && (i == 0 || i == nx - 2 || r == 0 || r == nz - 2);  // :312-315
      hy = fma(dbh, b + a, edge ? T(0) : hy_o);
    }
    return hy;
  };
  // Ex, Ez of one step at row r (:273-295), source, Dirichlet: the new Hy of
  // rows r-1, r (column i-1 for Ez)
  auto efield = [&](int r, T hy, T hy_p, T ex_c, T ez_c, const CE& k, MX& m, T ix, T iz, T& ex, T& ez) {
    const T hy_w = dpp_prev(hy);  // new Hy(i-1, r)
    const bool rv = valid(r);
    const int rs = crow(r);
    const bool inz = PML && in_slab(rs, nz, lz);
    ex = ex_c;
    ez = ez_c;
    if (INT || (rv && hcol && r >= 1 && r <= nz - 2)) {
      T d = (hy - hy_p) * idz;
      if (PML && inz) {
        m.hyz = fma(bzh[rs], m.hyz, azh[rs] * d);
        d = fma(d, ikzh[rs], m.hyz);
      }
      ex = fma(k.cax, ex, k.cbx * d);
    }
    if (INT || (rv && ezcol && r <= nz - 2)) {
      T d = (hy - hy_w) * idx;
      if (PML && inx) {
        m.hyx = fma(bxh, m.hyx, axh * d);
        d = fma(d, ikxh, m.hyx);
      }
      ez = fma(k.caz, ez, k.cbz * d);
    }
    if (r == ksrc && is_src_col) { ex += ix; ez += iz; }
    if (!INT) {
      if (i == 0 || i == nx - 2 || r == 0 || r == nz - 1) ex = 0;
      if (i == 0 || i == nx - 1 || r == 0 || r == nz - 2) ez = 0;
      if (i == nx - 1) ex = 0;  // outside the (nx-1) x nz Ex array
      if (r == nz - 1) ez = 0;  // outside the nx x (nz-1) Ez array
      if (!rv) { ex = 0; ez = 0; }
    }
  };
  const int nrec = S.nrec;
  auto gather = [&](int b, int itn, T ex, T ez) { gather_row<T>(S, b, itn, lane, xo - 2, 2, 1 + EW2E, ex, ez); };

  T ex0_c = valid(zo - 2) ? exi[cell(zo - 2)] : T(0);  // Ex of step it-1, row r
  T hy1_p = 0;                                          // Hy of step it, row r-1
  T ex1_p = 0, ez1_p = 0;                               // E of step it, row r-1
  T hy2_p = 0;                                          // Hy of step it+1, row r-2
  CE k_p{};                                             // coefficients of row r-1
  MX m_p{};                                             // memory of row r-1 after step it
  const bool has_rec = nrec > 0 && S.rrow[ze] > S.rrow[zo];  // receivers on own rows (per tile)
  bool bad0 = false, bad1 = false;

  // every load of row r (addresses clamped in bounds)
  struct Row { T exn, ez, hy; CE k; MX m; };
  auto load_row = [&](int r) {
    Row L;
    const uint32_t c = cell(r);
    L.exn = exi[cell(r + 1)];
    L.ez = ezi[c];
    L.hy = hyi[c];
    L.k.cax = CAX[c];
    L.k.cbx = CBX[c];
    L.k.caz = CAZ[c];
    L.k.cbz = CBZ[c];
    L.m = load_m(r);
    return L;
  };
  // row step r: A, B of step it at row r, then C, D of step it+1 at row r-1
  auto step_row = [&](const int r, const Row& L) {
    const bool actB = r >= zo - 1, actC = r >= zo, actD = r >= zo + 1;
    const T ld_exn = L.exn, ld_ez = L.ez, ld_hy = L.hy;
    const CE k = L.k;
    MX m = L.m;
    const bool rv = valid(r);
    const T ex0_n = valid(r + 1) ? ld_exn : T(0);
    const T ez0_c = rv ? ld_ez : T(0);
    const T hy0 = rv ? ld_hy : T(0);

    // A, B: step it at row r
    const T hy1_c = hfield(r, ex0_c, ex0_n, ez0_c, hy0, m);
    T ex1_c = 0, ez1_c = 0;
    if (actB) {
      efield(r, hy1_c, hy1_p, ex0_c, ez0_c, k, m, incx0, incz0, ex1_c, ez1_c);
      if (r >= zo && r < ze) {
        bad0 |= own && !((double)ex1_c * ex1_c + (double)ez1_c * ez1_c <= 1.0e50);
        if (has_rec) gather(r, it, ex1_c, ez1_c);
      }
    }
    // C, D: step it+1 at row r-1
    T hy2_c = 0;
    if (actC) {
      MX mc = m_p;
      hy2_c = hfield(r - 1, ex1_p, ex1_c, ez1_p, hy1_p, mc);
      if (actD && r - 1 < ze) {
        T ex2, ez2;
        efield(r - 1, hy2_c, hy2_p, ex1_p, ez1_p, k_p, mc, incx1, incz1, ex2, ez2);
        bad1 |= own && !((double)ex2 * ex2 + (double)ez2 * ez2 <= 1.0e50);
        if (own) {
          const uint32_t co = cell(r - 1);
          exo[co] = ex2;
          ezo[co] = ez2;
          hyo[co] = hy2_c;
          store_m(r - 1, mc);
        }
        if (has_rec) gather(r - 1, it + 1, ex2, ez2);
      }
    }
    ex0_c = ex0_n;
    hy1_p = hy1_c;
    ex1_p = ex1_c;
    ez1_p = ez1_c;
    hy2_p = hy2_c;
    k_p = k;
    m_p = m;
  };
  for (int r = zo - 2; r <= ze; ++r) step_row(r, load_row(r));
  if (bad0)
    check_blowup(HUGE_VAL, S.blowup, it);
  else if (bad1)
    check_blowup(HUGE_VAL, S.blowup, it + 1);
}

// Work: as k_em2_step, with EW2E owned columns per wavefront.
template <typename T, int WB>
__global__ void __launch_bounds__(64 * WB)
k_em2_step2(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const int4* __restrict__ work, int nitems, const int* __restrict__ base, int rel,
            const int* __restrict__ it0, int eh0) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int it = __builtin_amdgcn_readfirstlane(*base + rel);
  const int q = __builtin_amdgcn_readfirstlane(((it - *it0) >> 1) & 1);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    const int4 wi = make_int4(__builtin_amdgcn_readfirstlane(wv.x),
                              __builtin_amdgcn_readfirstlane(wv.y) + wave * EW2E,
                              __builtin_amdgcn_readfirstlane(wv.z), 0);
    if (wi.x < 0) continue;
    const DModel& M = models[shots[wi.x].model];
    if (wi.y >= M.nx) continue;
    const int eh = wv.w > 0 ? __builtin_amdgcn_readfirstlane(wv.w) : eh0;  // rows of this tile
    // every column (lanes) and row (zo-2 .. ze+1) the wave reads outside the
    // C-PML slabs and clear of the columns and rows the edge rules name
    const int ze = min(wi.z + eh, M.nz);
    const bool interior = wi.y - 2 >= M.lx + 1 && wi.y + 61 <= M.nx - M.lx - 3 &&
                          wi.z - 2 >= M.lz + 1 && ze + 1 <= M.nz - M.lz - 3;
    if (interior)
      em2x2_tile<T, false>(models, shots, wi, it, q, threadIdx.x & 63, eh);
    else
      em2x2_tile<T, true>(models, shots, wi, it, q, threadIdx.x & 63, eh);
  }
}

// Hy zeroed on its Dirichlet ring (:312-315) in buffer A at the end of a run,
// so the returned state equals the legacy state.
template <typename T>
__global__ void k_em2_edges(const DModel* __restrict__ models, const DShot* __restrict__ shots) {
  const DShot& S = shots[blockIdx.y];
  const DModel& M = models[S.model];
  const int nx = M.nx, nz = M.nz;
  auto hy = mp<T>(S.f[2]);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * (nx + nz); q += gridDim.x * blockDim.x) {
    int i, j;
    if (q < nx) { i = q; j = 0; }
    else if (q < 2 * nx) { i = q - nx; j = nz - 2; }
    else if (q < 2 * nx + nz) { i = 0; j = q - 2 * nx; }
    else { i = nx - 2; j = q - 2 * nx - nz; }
    if (i >= 0 && j >= 0) hy[(long)j * nx + i] = T(0);
  }
}

// ===========================================================================
// ELASTIC25: seismic25, legacy/fortran/seismicfdtd.f95:388-949
// Two launches per step: stresses, then velocities.  A wavefront takes one
// z plane k, 64 adjacent columns (lanes 1..EW25 are owned; lanes 0 and 63
// only supply x neighbours, which travel by lane shuffles) and a chunk of at
// most Y25 rows in y, which it walks with its y neighbours kept in registers.
// The 2-D (y-invariant) coefficients of its columns are loaded once per
// chunk; z neighbours are plain loads of the adjacent planes.  A block is
// XW25 x KB25 wavefronts (columns x planes).  Chunks end at the y-slab edges,
// and a wavefront clear of every slab runs a body without memory variables.
// psi slabs (per shot):
//   x [nz][ny][2lx]: dvx_dx dvy_dx dvz_dx dsxx_dx dsxy_dx dsxz_dx
//   y [nz][2ly][nx]: dvx_dy dvy_dy dvz_dy dsyy_dy dsxy_dy dsyz_dy
//   z [2lz][ny][nx]: dvx_dz dvy_dz dvz_dz dszz_dz dsxz_dz dsyz_dz
// ===========================================================================
#ifndef SDT_EL25_PLANES
#define SDT_EL25_PLANES 1
#endif
#ifndef SDT_EL25_ROWS
#define SDT_EL25_ROWS 32
#endif
#ifndef SDT_EL25_XW
#define SDT_EL25_XW 1
#endif
// (56, which puts every row of stores on a 32-byte boundary, measured slower
// here: 1.319 against 1.309 ms per aniso25 step, 0.640 against 0.593 ms per
// gpr25 step; the lanes lost cost more than the aligned stores gain)
#ifndef SDT_EL25_EW
#define SDT_EL25_EW 62
#endif
constexpr int EW25 = SDT_EL25_EW;      // owned columns per wavefront
static_assert(EW25 >= 1 && EW25 <= 62, "2.5-D wavefront: one halo lane on each side");
constexpr int KB25 = SDT_EL25_PLANES;  // z planes per block
constexpr int XW25 = SDT_EL25_XW;      // wavefronts side by side in x per block
constexpr int NT25 = 64 * KB25 * XW25;
constexpr int Y25 = SDT_EL25_ROWS;     // y rows per work item

// The y chunk [j0, j1) of a 2.5-D work item travels in one int as two
// unsigned 16-bit halves (ny <= 65535 is checked at plan creation); the
// decode shifts as unsigned so chunks ending at or above 32768 stay positive.
__host__ __device__ inline int ychunk_pack(int j0, int j1) {
  return (int)((unsigned)j0 | ((unsigned)j1 << 16));
}
__host__ __device__ inline int ychunk_j0(int w) { return (int)((unsigned)w & 0xffffu); }
__host__ __device__ inline int ychunk_j1(int w) { return (int)((unsigned)w >> 16); }

struct Slab3 {
  long SX, SY, SZ;       // elements per memory-variable array on each axis
  long bx, by, bz;       // base offsets of the x, y, z groups
  __device__ Slab3(int nx, int ny, int nz, int lx, int ly, int lz) {
    SX = (long)nz * ny * 2 * lx;
    SY = (long)nz * 2 * ly * nx;
    SZ = (long)2 * lz * ny * nx;
    bx = 0;
    by = 6 * SX;
    bz = 6 * SX + 6 * SY;
  }
};

// One z plane of a 3-D field: wave-uniform base, 32-bit offsets j * nx + i.
template <typename T>
__device__ __forceinline__ A32<T> plane(void* f, long nxy, int k) {
  return A32<T>(static_cast<T*>(f) + (long)k * nxy);
}
template <typename T>
__device__ __forceinline__ A32<const T> cplane(const void* f, long nxy, int k) {
  return A32<const T>(static_cast<const T*>(f) + (long)k * nxy);
}

// waves per SIMD the register budget is sized for (FP32; FP64 takes 2)
#ifndef SDT_EL25S_WAVES
#define SDT_EL25S_WAVES 3
#endif
#ifndef SDT_EL25V_WAVES
#define SDT_EL25V_WAVES 4
#endif
// the C-PML body of the stress pass with two rows of loads in flight too
// (measured no faster: 1.312 against 1.310 ms per FP32 aniso25 step, at more
// registers)
#ifndef SDT_EL25_PF_PML
#define SDT_EL25_PF_PML 0
#endif

template <typename T, bool PML>
__device__ __forceinline__ void el25_stress_tile(const DModel* __restrict__ models,
                                              const DShot* __restrict__ shots, const int4 wv,
                                              const int it, T* __restrict__ cs) {
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const int nx = M.nx, ny = M.ny, nz = M.nz;
  // a wavefront past the last plane stays (clamped, storing nothing) so that
  // every wavefront of the block takes the same barriers
  const int wx_ = (int)(threadIdx.x >> 6) % XW25;  // x slot in the block
  const int wv_ = (int)(threadIdx.x >> 6) / XW25;  // plane slot in the block
  const int kw = __builtin_amdgcn_readfirstlane(wv.z) + wv_;
  const bool live = kw < nz;
  const int k = live ? kw : nz - 1;
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(wv.y) + wx_ * EW25 - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);  // clamped column: loads stay in bounds
  const bool own = live && incol && lane >= 1 && lane <= EW25;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  const long nxy = (long)nx * ny;
  const uint32_t p = (uint32_t)k * nx + ic;
  const T idx = T(1.0 / M.dx), idy = T(1.0 / M.dy), idz = T(1.0 / M.dz), dt = T(M.dt);
  const int lx = M.lx, ly = M.ly, lz = M.lz;
  const bool inx = PML && incol && in_slab(i, nx, lx), inz = PML && in_slab(k, nz, lz);
  const Prof<T> P(&M.prof[0][0]);
  gcptr<T> iky = P[6 + PK], ay = P[6 + PA], by = P[6 + PB];
  gcptr<T> ikyh = P[6 + PKH], ayh = P[6 + PAH], byh = P[6 + PBH];
  const T bx_i = P[PB][ic], ax_i = P[PA][ic], ikx_i = P[PK][ic];
  const T bxh_i = P[PBH][ic], axh_i = P[PAH][ic], ikxh_i = P[PKH][ic];
  const T bz_k = P[12 + PB][k], az_k = P[12 + PA][k], ikz_k = P[12 + PK][k];
  const T bzh_k = P[12 + PBH][k], azh_k = P[12 + PAH][k], ikzh_k = P[12 + PKH][k];
  // loop ranges of the four legacy stress loops (seismicfdtd.f95:617, :677, :728, :771)
  const bool p1 = own && k >= 1 && i <= nx - 2, p2 = own && k >= 1 && i >= 1;
  const bool p3 = own && k <= nz - 2 && i >= 1, p4 = own && k <= nz - 2 && i <= nx - 2;
  // The 33 coefficients of the four loops at this column and plane do not
  // change along y: they are loaded once per work item into LDS, one slot per
  // thread that only that thread reads back, and read where each loop uses
  // them (volatile, so they are not held in registers for the whole walk);
  // the registers go to a second set of row loads instead.  cs: the block's
  // 33 * NT25 LDS slots.
  {
    T* w = cs + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 15; ++q) w[q * NT25] = p1 ? A32<const T>(M.coef[q])[p] : T(0);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      w[(15 + q) * NT25] = p2 ? A32<const T>(M.coef[15 + q])[p] : T(0);
      w[(21 + q) * NT25] = p3 ? A32<const T>(M.coef[21 + q])[p] : T(0);
      w[(27 + q) * NT25] = p4 ? A32<const T>(M.coef[27 + q])[p] : T(0);
    }
  }
  // coefficient q at C[q * NT25], read through an LDS pointer (ds_read)
  const __attribute__((address_space(3))) volatile T* C =
      (const __attribute__((address_space(3))) volatile T*)(cs) + threadIdx.x;
  const int kd = k >= 1 ? k - 1 : 0, ku = k <= nz - 2 ? k + 1 : nz - 1;
  const A32<const T> VX = cplane<T>(S.f[0], nxy, k), VY = cplane<T>(S.f[1], nxy, k);
  const A32<const T> VZ = cplane<T>(S.f[2], nxy, k);
  const A32<const T> VXD = cplane<T>(S.f[0], nxy, kd), VYD = cplane<T>(S.f[1], nxy, kd);
  const A32<const T> VZD = cplane<T>(S.f[2], nxy, kd);
  const A32<const T> VXU = cplane<T>(S.f[0], nxy, ku), VYU = cplane<T>(S.f[1], nxy, ku);
  const A32<const T> VZU = cplane<T>(S.f[2], nxy, ku);
  const A32<T> SXX = plane<T>(S.f[3], nxy, k), SYY = plane<T>(S.f[4], nxy, k);
  const A32<T> SZZ = plane<T>(S.f[5], nxy, k), SXY = plane<T>(S.f[6], nxy, k);
  const A32<T> SXZ = plane<T>(S.f[7], nxy, k), SYZ = plane<T>(S.f[8], nxy, k);
  auto psi = mp<T>(S.psi);
  const Slab3 L(nx, ny, nz, lx, ly, lz);
  const int xs = inx ? slab_pos(i, nx, lx) : 0, zs = inz ? slab_pos(k, nz, lz) : 0;

  // y window of the velocities: rows j-1 (m), j (c)
  const uint32_t r0 = (uint32_t)j0 * nx + ic;
  T vxm = 0, vym = 0, vzm = 0;
  if (j0 >= 1) { vxm = VX[r0 - nx]; vym = VY[r0 - nx]; vzm = VZ[r0 - nx]; }
  T vxc = VX[r0], vyc = VY[r0], vzc = VZ[r0];

  // the loads of row j: velocities of row j+1 (n) and of planes k-1 (d) and
  // k+1 (u), the row's stresses (rows past the last are clamped in bounds)
  struct RowLd { T vxn, vyn, vzn, vxd, vyd, vzd, vxu, vyu, vzu, sxx, syy, szz, sxy, sxz, syz; };
  auto load_row = [&](int j) {
    RowLd R;
    const uint32_t c = (uint32_t)(j < ny ? j : ny - 1) * nx + ic;
    const uint32_t cn = j + 1 < ny ? c + nx : c;
    R.vxn = VX[cn]; R.vyn = VY[cn]; R.vzn = VZ[cn];
    R.vxd = VXD[c]; R.vyd = VYD[c]; R.vzd = VZD[c];
    R.vxu = VXU[c]; R.vyu = VYU[c]; R.vzu = VZU[c];
    R.sxx = SXX[c]; R.syy = SYY[c]; R.szz = SZZ[c];
    R.sxy = SXY[c]; R.sxz = SXZ[c]; R.syz = SYZ[c];
    return R;
  };
  auto row = [&](const int j, const RowLd& R) {
    const uint32_t c = (uint32_t)j * nx + ic;
    const T vxn = R.vxn, vyn = R.vyn, vzn = R.vzn;      // row j+1 (n)
    const T vxd = R.vxd, vyd = R.vyd, vzd = R.vzd;      // plane k-1 (d)
    const T vxu = R.vxu, vyu = R.vyu, vzu = R.vzu;      // plane k+1 (u)
    const T sxx0 = R.sxx, syy0 = R.syy, szz0 = R.szz, sxy0 = R.sxy, sxz0 = R.sxz, syz0 = R.syz;
    const T vxe = lane_next(vxc), vye = lane_next(vyc), vze = lane_next(vzc);  // i+1 (e)
    const T vxw = lane_prev(vxc), vyw = lane_prev(vyc), vzw = lane_prev(vzc);  // i-1 (w)
    const bool iny = PML && in_slab(j, ny, ly);
    const long ox = L.bx + ((long)k * ny + j) * 2 * lx + xs;
    const long oy = L.by + ((long)k * 2 * ly + (iny ? slab_pos(j, ny, ly) : 0)) * nx + ic;
    const long oz = L.bz + ((long)zs * ny + j) * nx + ic;
    // velocity-derivative memory variables (shared by all four loops: A1)
    T mxx = 0, myx = 0, mzx = 0, mxy = 0, myy = 0, mzy = 0, mxz = 0, myz = 0, mzz = 0;
    if (inx) { mxx = psi[ox]; myx = psi[ox + L.SX]; mzx = psi[ox + 2 * L.SX]; }
    if (iny && incol) { mxy = psi[oy]; myy = psi[oy + L.SY]; mzy = psi[oy + 2 * L.SY]; }
    if (inz && incol) { mxz = psi[oz]; myz = psi[oz + L.SZ]; mzz = psi[oz + 2 * L.SZ]; }

    // loop 1: sigma xx, yy, zz (seismicfdtd.f95:617-674)
    if (p1 && j >= 1) {
      T dvx_dx = (vxe - vxc) * idx, dvy_dx = (vye - vyc) * idx, dvz_dx = (vze - vzc) * idx;
      T dvy_dy = (vyc - vym) * idy, dvx_dy = (vxc - vxm) * idy, dvz_dy = (vzc - vzm) * idy;
      T dvz_dz = (vzc - vzd) * idz, dvx_dz = (vxc - vxd) * idz, dvy_dz = (vyc - vyd) * idz;
      if (inx) {
        mxx = bxh_i * mxx + axh_i * dvx_dx; myx = bxh_i * myx + axh_i * dvy_dx; mzx = bxh_i * mzx + axh_i * dvz_dx;
        dvx_dx = dvx_dx * ikxh_i + mxx; dvy_dx = dvy_dx * ikxh_i + myx; dvz_dx = dvz_dx * ikxh_i + mzx;
      }
      if (iny) {
        const T b = by[j], a = ay[j], ik = iky[j];
        myy = b * myy + a * dvy_dy; mxy = b * mxy + a * dvx_dy; mzy = b * mzy + a * dvz_dy;
        dvy_dy = dvy_dy * ik + myy; dvx_dy = dvx_dy * ik + mxy; dvz_dy = dvz_dy * ik + mzy;
      }
      if (inz) {
        mzz = bz_k * mzz + az_k * dvz_dz; mxz = bz_k * mxz + az_k * dvx_dz; myz = bz_k * myz + az_k * dvy_dz;
        dvz_dz = dvz_dz * ikz_k + mzz; dvx_dz = dvx_dz * ikz_k + mxz; dvy_dz = dvy_dz * ikz_k + myz;
      }
      const T c0 = C[0], c1 = C[1 * NT25], c2 = C[2 * NT25], c3 = C[3 * NT25], c4 = C[4 * NT25];
      const T c5 = C[5 * NT25], c6 = C[6 * NT25], c7 = C[7 * NT25], c8 = C[8 * NT25], c9 = C[9 * NT25];
      const T c10 = C[10 * NT25], c11 = C[11 * NT25], c12 = C[12 * NT25], c13 = C[13 * NT25];
      const T c14 = C[14 * NT25];
      SXX[c] = sxx0 + (c0 * dvx_dx + c1 * dvy_dy + c2 * dvz_dz + c3 * (dvy_dz + dvz_dy) +
                       c4 * (dvx_dz + dvz_dx) + c5 * (dvx_dy + dvz_dy)) * dt;  // A3 (c16)
      SYY[c] = syy0 + (c1 * dvx_dx + c6 * dvy_dy + c7 * dvz_dz + c8 * (dvy_dz + dvz_dy) +
                       c9 * (dvx_dz + dvz_dx) + c10 * (dvy_dx + dvx_dy)) * dt;
      SZZ[c] = szz0 + (c2 * dvx_dx + c7 * dvy_dy + c11 * dvz_dz + c12 * (dvy_dz + dvz_dy) +
                       c13 * (dvx_dz + dvz_dx) + c14 * (dvy_dx + dvx_dy)) * dt;
    }
    // loop 2: sigma xy (seismicfdtd.f95:677-725); dvz_dx and dvx_dz get no correction
    if (p2 && j <= ny - 2) {
      T dvx_dx = (vxc - vxw) * idx, dvy_dx = (vyc - vyw) * idx, dvz_dx = (vzc - vzw) * idx;
      T dvy_dy = (vyn - vyc) * idy, dvx_dy = (vxn - vxc) * idy, dvz_dy = (vzn - vzc) * idy;
      T dvz_dz = (vzc - vzd) * idz, dvx_dz = (vxc - vxd) * idz, dvy_dz = (vyc - vyd) * idz;
      if (inx) {
        mxx = bx_i * mxx + ax_i * dvx_dx; myx = bx_i * myx + ax_i * dvy_dx; mzx = bx_i * mzx + ax_i * dvz_dx;
        dvx_dx = dvx_dx * ikx_i + mxx; dvy_dx = dvy_dx * ikx_i + myx;
      }
      if (iny) {
        const T b = byh[j], a = ayh[j], ik = ikyh[j];
        myy = b * myy + a * dvy_dy; mxy = b * mxy + a * dvx_dy; mzy = b * mzy + a * dvz_dy;
        dvy_dy = dvy_dy * ik + myy; dvx_dy = dvx_dy * ik + mxy; dvz_dy = dvz_dy * ik + mzy;
      }
      if (inz) {
        mzz = bzh_k * mzz + azh_k * dvz_dz; mxz = bzh_k * mxz + azh_k * dvx_dz; myz = bzh_k * myz + azh_k * dvy_dz;
        dvz_dz = dvz_dz * ikzh_k + mzz; dvy_dz = dvy_dz * ikzh_k + myz;
      }
      const __attribute__((address_space(3))) volatile T* K = C + 15 * NT25;
      SXY[c] = sxy0 + (K[0] * dvx_dx + K[NT25] * dvy_dy + K[2 * NT25] * dvz_dz + K[3 * NT25] * (dvz_dy + dvy_dz) +
                       K[4 * NT25] * (dvz_dx + dvx_dz) + K[5 * NT25] * (dvy_dx + dvx_dy)) * dt;
    }
    // loop 3: sigma xz (seismicfdtd.f95:728-768)
    if (p3 && j >= 1) {
      T dvx_dx = (vxc - vxw) * idx, dvy_dx = (vyc - vyw) * idx, dvz_dx = (vzc - vzw) * idx;
      T dvy_dy = (vyc - vym) * idy, dvz_dy = (vzc - vzm) * idy, dvx_dy = (vxc - vxm) * idy;
      T dvz_dz = (vzu - vzc) * idz, dvx_dz = (vxu - vxc) * idz, dvy_dz = (vyu - vyc) * idz;
      if (inx) {
        mxx = bx_i * mxx + ax_i * dvx_dx; myx = bx_i * myx + ax_i * dvy_dx; mzx = bx_i * mzx + ax_i * dvz_dx;
        dvx_dx = dvx_dx * ikx_i + mxx; dvy_dx = dvy_dx * ikx_i + myx; dvz_dx = dvz_dx * ikx_i + mzx;
      }
      if (iny) {
        const T b = by[j], a = ay[j], ik = iky[j];
        myy = b * myy + a * dvy_dy; mxy = b * mxy + a * dvx_dy; mzy = b * mzy + a * dvz_dy;
        dvy_dy = dvy_dy * ik + myy; dvx_dy = dvx_dy * ik + mxy; dvz_dy = dvz_dy * ik + mzy;
      }
      if (inz) {
        mzz = bzh_k * mzz + azh_k * dvz_dz; mxz = bzh_k * mxz + azh_k * dvx_dz; myz = bzh_k * myz + azh_k * dvy_dz;
        dvz_dz = dvz_dz * ikzh_k + mzz; dvx_dz = dvx_dz * ikzh_k + mxz; dvy_dz = dvy_dz * ikzh_k + myz;
      }
      const __attribute__((address_space(3))) volatile T* K = C + 21 * NT25;
      SXZ[c] = sxz0 + (K[0] * dvx_dx + K[NT25] * dvy_dy + K[2 * NT25] * dvz_dz + K[3 * NT25] * (dvx_dz + dvz_dx) +
                       K[4 * NT25] * (dvx_dz + dvz_dx) + K[5 * NT25] * (dvx_dy + dvy_dx)) * dt;  // A3 (c45)
    }
    // loop 4: sigma yz (seismicfdtd.f95:771-813)
    if (p4 && j <= ny - 2) {
      T dvx_dx = (vxe - vxc) * idx, dvy_dx = (vye - vyc) * idx, dvz_dx = (vze - vzc) * idx;
      T dvy_dy = (vyn - vyc) * idy, dvx_dy = (vxn - vxc) * idy, dvz_dy = (vzn - vzc) * idy;
      T dvz_dz = (vzu - vzc) * idz, dvx_dz = (vxu - vxc) * idz, dvy_dz = (vyu - vyc) * idz;
      if (inx) {
        mxx = bxh_i * mxx + axh_i * dvx_dx; myx = bxh_i * myx + axh_i * dvy_dx; mzx = bxh_i * mzx + axh_i * dvz_dx;
        dvx_dx = dvx_dx * ikxh_i + mxx; dvy_dx = dvy_dx * ikxh_i + myx; dvz_dx = dvz_dx * ikxh_i + mzx;
      }
      if (iny) {
        const T b = byh[j], a = ayh[j], ik = ikyh[j];
        myy = b * myy + a * dvy_dy; mxy = b * mxy + a * dvx_dy; mzy = b * mzy + a * dvz_dy;
        dvy_dy = dvy_dy * ik + myy; dvx_dy = dvx_dy * ik + mxy; dvz_dy = dvz_dy * ik + mzy;
      }
      if (inz) {
        mzz = bzh_k * mzz + azh_k * dvz_dz; mxz = bzh_k * mxz + azh_k * dvx_dz; myz = bzh_k * myz + azh_k * dvy_dz;
        dvz_dz = dvz_dz * ikzh_k + mzz; dvx_dz = dvx_dz * ikzh_k + mxz; dvy_dz = dvy_dz * ikzh_k + myz;
      }
      const __attribute__((address_space(3))) volatile T* K = C + 27 * NT25;
      SYZ[c] = syz0 + (K[0] * dvx_dx + K[NT25] * dvy_dy + K[2 * NT25] * dvz_dz + K[3 * NT25] * (dvy_dz + dvz_dy) +
                       K[4 * NT25] * (dvx_dz + dvz_dx) + K[5 * NT25] * (dvy_dx + dvx_dy)) * dt;
    }
    if (own) {
      if (inx) { psi[ox] = mxx; psi[ox + L.SX] = myx; psi[ox + 2 * L.SX] = mzx; }
      if (iny) { psi[oy] = mxy; psi[oy + L.SY] = myy; psi[oy + 2 * L.SY] = mzy; }
      if (inz) { psi[oz] = mxz; psi[oz + L.SZ] = myz; psi[oz + 2 * L.SZ] = mzz; }
      if (i == S.isrc && k == S.ksrc && j == S.jsrc) {  // stress-rate (moment-tensor) source
        if (S.msrc[0]) SXX[c] = SXX[c] + cp<T>(S.msrc[0])[it];
        if (S.msrc[1]) SYY[c] = SYY[c] + cp<T>(S.msrc[1])[it];
        if (S.msrc[2]) SZZ[c] = SZZ[c] + cp<T>(S.msrc[2])[it];
        if (S.msrc[3]) SXY[c] = SXY[c] + cp<T>(S.msrc[3])[it];
        if (S.msrc[4]) SXZ[c] = SXZ[c] + cp<T>(S.msrc[4])[it];
        if (S.msrc[5]) SYZ[c] = SYZ[c] + cp<T>(S.msrc[5])[it];
      }
    }
    vxm = vxc; vym = vyc; vzm = vzc;
    vxc = vxn; vyc = vyn; vzc = vzn;
  };
  // two rows per iteration: the loads of row j+1 are in flight while row j
  // computes, and those of row j+2 while row j+1 does
  if constexpr (!PML || SDT_EL25_PF_PML) {
    RowLd A = load_row(j0);
    int j = j0;
#pragma unroll 1
    for (; j + 1 < j1; j += 2) {
      const RowLd B = load_row(j + 1);
      row(j, A);
      A = load_row(j + 2);
      row(j + 1, B);
    }
    if (j < j1) row(j, A);
  } else {
#pragma unroll 1
    for (int j = j0; j < j1; ++j) row(j, load_row(j));
  }
}

// A wavefront whose columns, plane and rows all lie outside the C-PML slabs
// runs the body without memory variables.
template <typename T>
__global__ void __launch_bounds__(NT25, sizeof(T) == 4 ? SDT_EL25S_WAVES : 2)
k_el25_stress(const DModel* __restrict__ models, const DShot* __restrict__ shots,
          const int4* __restrict__ work, const int* __restrict__ base, int rel) {
  const int4 wv = work[blockIdx.x];
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  if (s < 0) return;
  const DModel& M = models[shots[s].model];
  const int xo = __builtin_amdgcn_readfirstlane(wv.y) + (int)((threadIdx.x >> 6) % XW25) * EW25;
  const int k = __builtin_amdgcn_readfirstlane(wv.z) + (int)(threadIdx.x >> 6) / XW25;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  const bool interior = xo - 1 >= M.lx && xo + 62 < M.nx - M.lx && k >= M.lz && k < M.nz - M.lz &&
                        j0 >= M.ly && j1 <= M.ny - M.ly;
  const int it = *base + rel;
  __shared__ T cs[33 * NT25];  // the y-invariant coefficients (el25_stress_tile)
  if (interior)
    el25_stress_tile<T, false>(models, shots, wv, it, cs);
  else
    el25_stress_tile<T, true>(models, shots, wv, it, cs);
}

template <typename T, bool PML>
__device__ __forceinline__ void el25_vel_tile(const DModel* __restrict__ models,
                                              const DShot* __restrict__ shots, const int4 wv, const int* __restrict__ base, int rel) {
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const int nx = M.nx, ny = M.ny, nz = M.nz;
  // a wavefront past the last plane stays (clamped, storing nothing) so that
  // every wavefront of the block takes the same barriers
  const int wx_ = (int)(threadIdx.x >> 6) % XW25;  // x slot in the block
  const int wv_ = (int)(threadIdx.x >> 6) / XW25;  // plane slot in the block
  const int kw = __builtin_amdgcn_readfirstlane(wv.z) + wv_;
  const bool live = kw < nz;
  const int k = live ? kw : nz - 1;
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(wv.y) + wx_ * EW25 - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);
  const bool own = live && incol && lane >= 1 && lane <= EW25;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  const int it = *base + rel;
  const long nxy = (long)nx * ny;
  const uint32_t p = (uint32_t)k * nx + ic;
  const T idx = T(1.0 / M.dx), idy = T(1.0 / M.dy), idz = T(1.0 / M.dz);
  const int lx = M.lx, ly = M.ly, lz = M.lz;
  const bool inx = PML && incol && in_slab(i, nx, lx), inz = PML && in_slab(k, nz, lz);
  const Prof<T> P(&M.prof[0][0]);
  const T bx_i = P[PB][ic], ax_i = P[PA][ic], ikx_i = P[PK][ic];
  const T bxh_i = P[PBH][ic], axh_i = P[PAH][ic], ikxh_i = P[PKH][ic];
  const T bz_k = P[12 + PB][k], az_k = P[12 + PA][k], ikz_k = P[12 + PK][k];
  const T bzh_k = P[12 + PBH][k], azh_k = P[12 + PAH][k], ikzh_k = P[12 + PKH][k];
  gcptr<T> by = P[6 + PB], ay = P[6 + PA], iky = P[6 + PK];
  gcptr<T> byh = P[6 + PBH], ayh = P[6 + PAH], ikyh = P[6 + PKH];
  // loop ranges of the legacy velocity loops (seismicfdtd.f95:821, :848, :873)
  const bool ux = own && k >= 1 && i >= 1, uy = own && k >= 1 && i <= nx - 2;
  const bool uz = own && k <= nz - 2 && i <= nx - 2;
  const T dtrx = ux ? A32<const T>(M.coef[33])[p] : T(0), dtry = uy ? A32<const T>(M.coef[34])[p] : T(0);
  const T dtrz = uz ? A32<const T>(M.coef[35])[p] : T(0);
  const T omx = A32<const T>(M.coef[36])[p], omy = A32<const T>(M.coef[37])[p];
  const T omz = A32<const T>(M.coef[38])[p];
  const int kd = k >= 1 ? k - 1 : 0, ku = k <= nz - 2 ? k + 1 : nz - 1;
  const A32<const T> SXX = cplane<T>(S.f[3], nxy, k), SYY = cplane<T>(S.f[4], nxy, k);
  const A32<const T> SZZ = cplane<T>(S.f[5], nxy, k), SXY = cplane<T>(S.f[6], nxy, k);
  const A32<const T> SXZ = cplane<T>(S.f[7], nxy, k), SYZ = cplane<T>(S.f[8], nxy, k);
  const A32<const T> SXZD = cplane<T>(S.f[7], nxy, kd), SYZD = cplane<T>(S.f[8], nxy, kd);
  const A32<const T> SZZU = cplane<T>(S.f[5], nxy, ku);
  const A32<T> VX = plane<T>(S.f[0], nxy, k), VY = plane<T>(S.f[1], nxy, k), VZ = plane<T>(S.f[2], nxy, k);
  auto psi = mp<T>(S.psi);
  const Slab3 L(nx, ny, nz, lx, ly, lz);
  const int xs = inx ? slab_pos(i, nx, lx) : 0, zs = inz ? slab_pos(k, nz, lz) : 0;
  const bool kedge = (i == 0 || k == 0 || i == nx - 1 || k == nz - 1);
  const bool srccol = (i == S.isrc && k == S.ksrc);

  // y window: sxy, syz of row j-1 and syy of row j
  const uint32_t r0 = (uint32_t)j0 * nx + ic;
  T sxym = 0, syzm = 0;
  if (j0 >= 1) { sxym = SXY[r0 - nx]; syzm = SYZ[r0 - nx]; }
  T syyc = SYY[r0];
  // the loads of row j (rows past the last are clamped in bounds)
  struct RowLd { T sxx, sxy, sxz, szz, syz, syyn, sxzd, syzd, szzu, vx, vy, vz; };
  auto load_row = [&](int j) {
    RowLd R;
    const uint32_t c = (uint32_t)(j < ny ? j : ny - 1) * nx + ic;
    const uint32_t cn = j + 1 < ny ? c + nx : c;
    R.sxx = SXX[c]; R.sxy = SXY[c]; R.sxz = SXZ[c]; R.szz = SZZ[c]; R.syz = SYZ[c];
    R.syyn = SYY[cn];
    R.sxzd = SXZD[c]; R.syzd = SYZD[c]; R.szzu = SZZU[c];
    R.vx = VX[c]; R.vy = VY[c]; R.vz = VZ[c];
    return R;
  };
  auto row = [&](const int j, const RowLd& R) {
    const uint32_t c = (uint32_t)j * nx + ic;
    const T sxx = R.sxx, sxy = R.sxy, sxz = R.sxz, szz = R.szz, syz = R.syz;
    const T syyn = R.syyn;
    const T sxzd = R.sxzd, syzd = R.syzd, szzu = R.szzu;
    const T sxxw = lane_prev(sxx), sxye = lane_next(sxy), sxze = lane_next(sxz);
    T vxn = R.vx, vyn = R.vy, vzn = R.vz;
    const bool iny = PML && in_slab(j, ny, ly);
    const long ox = L.bx + ((long)k * ny + j) * 2 * lx + xs;
    const long oy = L.by + ((long)k * 2 * ly + (iny ? slab_pos(j, ny, ly) : 0)) * nx + ic;
    const long oz = L.bz + ((long)zs * ny + j) * nx + ic;
    // vx, seismicfdtd.f95:821-846
    if (ux && j >= 1) {
      T a = (sxx - sxxw) * idx, b = (sxy - sxym) * idy, d = (sxz - sxzd) * idz;
      if (inx) { T m = bx_i * psi[ox + 3 * L.SX] + ax_i * a; a = a * ikx_i + m; psi[ox + 3 * L.SX] = m; }
      if (iny) { T m = by[j] * psi[oy + 4 * L.SY] + ay[j] * b; b = b * iky[j] + m; psi[oy + 4 * L.SY] = m; }
      if (inz) { T m = bz_k * psi[oz + 4 * L.SZ] + az_k * d; d = d * ikz_k + m; psi[oz + 4 * L.SZ] = m; }
      vxn = vxn * omx + (a + b + d) * dtrx;
    }
    // vy, seismicfdtd.f95:848-870
    if (uy && j <= ny - 2) {
      T a = (sxye - sxy) * idx, b = (syyn - syyc) * idy, d = (syz - syzd) * idz;
      if (inx) { T m = bxh_i * psi[ox + 4 * L.SX] + axh_i * a; a = a * ikxh_i + m; psi[ox + 4 * L.SX] = m; }
      if (iny) { T m = byh[j] * psi[oy + 3 * L.SY] + ayh[j] * b; b = b * ikyh[j] + m; psi[oy + 3 * L.SY] = m; }
      if (inz) { T m = bz_k * psi[oz + 5 * L.SZ] + az_k * d; d = d * ikz_k + m; psi[oz + 5 * L.SZ] = m; }
      vyn = vyn * omy + (a + b + d) * dtry;
    }
    // vz, seismicfdtd.f95:873-896
    if (uz && j >= 1) {
      T a = (sxze - sxz) * idx, b = (syz - syzm) * idy, d = (szzu - szz) * idz;
      if (inx) { T m = bxh_i * psi[ox + 5 * L.SX] + axh_i * a; a = a * ikxh_i + m; psi[ox + 5 * L.SX] = m; }
      if (iny) { T m = by[j] * psi[oy + 5 * L.SY] + ay[j] * b; b = b * iky[j] + m; psi[oy + 5 * L.SY] = m; }
      if (inz) { T m = bzh_k * psi[oz + 3 * L.SZ] + azh_k * d; d = d * ikzh_k + m; psi[oz + 3 * L.SZ] = m; }
      vzn = vzn * omz + (a + b + d) * dtrz;
    }
    // source, seismicfdtd.f95:898-900
    if (srccol && j == S.jsrc) {
      if (S.src[0]) vxn += cp<T>(S.src[0])[it];
      if (S.src[1]) vyn += cp<T>(S.src[1])[it];
      if (S.src[2]) vzn += cp<T>(S.src[2])[it];
    }
    // Dirichlet faces, seismicfdtd.f95:903-925
    if (kedge || j == 0 || j == ny - 1) { vxn = 0; vyn = 0; vzn = 0; }
    if (own) {
      check_blowup((double)vxn * vxn + (double)vyn * vyn + (double)vzn * vzn, S.blowup, it);
      VX[c] = vxn;
      VY[c] = vyn;
      VZ[c] = vzn;
    }
    sxym = sxy; syzm = syz; syyc = syyn;
  };
  // two rows per iteration, as in the stress pass (FP64: one row at a time,
  // two would cost a wavefront per SIMD)
  if constexpr (sizeof(T) == 4) {
    RowLd A = load_row(j0);
    int j = j0;
#pragma unroll 1
    for (; j + 1 < j1; j += 2) {
      const RowLd B = load_row(j + 1);
      row(j, A);
      A = load_row(j + 2);
      row(j + 1, B);
    }
    if (j < j1) row(j, A);
  } else {
#pragma unroll 1
    for (int j = j0; j < j1; ++j) row(j, load_row(j));
  }
}

// A wavefront whose columns, plane and rows all lie outside the C-PML slabs
// runs the body without memory variables.
template <typename T>
__global__ void __launch_bounds__(NT25, sizeof(T) == 4 ? SDT_EL25V_WAVES : 2)
k_el25_vel(const DModel* __restrict__ models, const DShot* __restrict__ shots,
          const int4* __restrict__ work, const int* __restrict__ base, int rel) {
  const int4 wv = work[blockIdx.x];
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  if (s < 0) return;
  const DModel& M = models[shots[s].model];
  const int xo = __builtin_amdgcn_readfirstlane(wv.y) + (int)((threadIdx.x >> 6) % XW25) * EW25;
  const int k = __builtin_amdgcn_readfirstlane(wv.z) + (int)(threadIdx.x >> 6) / XW25;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  const bool interior = xo - 1 >= M.lx && xo + 62 < M.nx - M.lx && k >= M.lz && k < M.nz - M.lz &&
                        j0 >= M.ly && j1 <= M.ny - M.ly;
  if (interior)
    el25_vel_tile<T, false>(models, shots, wv, base, rel);
  else
    el25_vel_tile<T, true>(models, shots, wv, base, rel);
}

// ===========================================================================
// EM2 adjoint: the transpose of one EM2 step (k_em2_step), for gradients
// of a trace misfit with respect to the E-update coefficients (caEx cbEx
// caEz cbEz, from permittivity and conductivity).  No counterpart in the
// reference.  The step as the device runs it: Hy' = m * Hy + dbH (dEz_dx +
// dEx_dz) with m = 0 on the Hy Dirichlet ring (the legacy zeroes it at the
// end of the previous step, electromagfdtd.f95:312-315), then Ex, Ez from
// Hy' (:273-295), source, Dirichlet (:304-311).  Three launches, like the
// elastic reference kernels: E^T (scratch planes 0, 1: adjoints of dHy_dz
// at Ex nodes and of dHy_dx at Ez nodes), H^T (gathers them into lambda_Hy;
// planes 2, 3: adjoints of dEx_dz and dEz_dx), and the E gather.
// psi slabs: x [nz][2lx]: dEz_dx dHy_dx ; z [2lz][nx]: dEx_dz dHy_dz.
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(NT)
k_em2_adj_e(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k,
            const int* __restrict__ tsel, unsigned gmask) {
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const int t = *tbase + k;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  const int lx = M.lx, lz = M.lz;
  const long SX = (long)nz * 2 * lx, SZ = (long)2 * lz * nx;
  const auto Lin = cp<T>(A.lam[(t + 1) & 1]);
  auto Lout = mp<T>(A.lam[t & 1]);
  const auto u1 = cp<T>(A.trail[*tsel]) + (long)(k + 1) * A.slot;
  const auto u0 = cp<T>(A.trail[*tsel]) + (long)k * A.slot;
  auto W = mp<T>(A.work);
  auto G = GAcc<T>::of(A);
  const Prof<T> P(&M.prof[0][0]);
  const bool inx = in_slab(i, nx, lx), inz = in_slab(j, nz, lz);
  const long ox = 3 * n + (inx ? (long)j * 2 * lx + slab_pos(i, nx, lx) : 0);
  const long oz = 3 * n + 2 * SX + (inz ? (long)slab_pos(j, nz, lz) * nx + i : 0);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz);
  // Dirichlet ring of Ex and Ez (and cells outside their Fortran extent)
  const bool edx = (i == 0 || i >= nx - 2 || j == 0 || j == nz - 1);
  const bool edz = (i == 0 || i == nx - 1 || j == 0 || j >= nz - 2);
  const T lex = edx ? T(0) : Lin[c], lez = edz ? T(0) : Lin[n + c];
  if (A.gsrc && i == S.isrc && j == S.ksrc) {
    A.gsrc[3 * (long)t + 0] = (double)lex;
    A.gsrc[3 * (long)t + 2] = (double)lez;
  }
  T wx = 0, wz = 0;
  if (j >= 1 && j <= nz - 2 && i <= nx - 2) {  // Ex node, :273-283
    const T ld = cp<T>(M.coef[1])[c] * lex;
    const T d = (u1[2 * n + c] - u1[2 * n + c - nx]) * idz;
    T D = d;
    wx = ld;
    if (inz) {
      const T lr = Lin[oz + SZ] + ld;
      wx = P[12 + PKH][j] * ld + P[12 + PAH][j] * lr;
      Lout[oz + SZ] = P[12 + PBH][j] * lr;
      D = P[12 + PKH][j] * d + u1[oz + SZ];
    }
    if (gmask & 1u) gadd(G + c, (double)lex * (double)u0[c]);
    if (gmask & 2u) gadd(G + n + c, (double)lex * (double)D);
    Lout[c] = cp<T>(M.coef[0])[c] * lex;
  } else {
    if (inz) Lout[oz + SZ] = Lin[oz + SZ];
    Lout[c] = lex;
  }
  if (j <= nz - 2 && i >= 1 && i <= nx - 2) {  // Ez node, :285-295
    const T le = cp<T>(M.coef[3])[c] * lez;
    const T e = (u1[2 * n + c] - u1[2 * n + c - 1]) * idx;
    T E = e;
    wz = le;
    if (inx) {
      const T lr = Lin[ox + SX] + le;
      wz = P[PKH][i] * le + P[PAH][i] * lr;
      Lout[ox + SX] = P[PBH][i] * lr;
      E = P[PKH][i] * e + u1[ox + SX];
    }
    if (gmask & 4u) gadd(G + 2 * n + c, (double)lez * (double)u0[n + c]);
    if (gmask & 8u) gadd(G + 3 * n + c, (double)lez * (double)E);
    Lout[n + c] = cp<T>(M.coef[2])[c] * lez;
  } else {
    if (inx) Lout[ox + SX] = Lin[ox + SX];
    Lout[n + c] = lez;
  }
  W[c] = wx;
  W[n + c] = wz;
}

template <typename T>
__global__ void __launch_bounds__(NT)
k_em2_adj_h(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k) {
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const int t = *tbase + k;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  const int lx = M.lx, lz = M.lz;
  const long SX = (long)nz * 2 * lx;
  const auto Lin = cp<T>(A.lam[(t + 1) & 1]);
  auto Lout = mp<T>(A.lam[t & 1]);
  auto W = mp<T>(A.work);
  const Prof<T> P(&M.prof[0][0]);
  const bool inx = in_slab(i, nx, lx), inz = in_slab(j, nz, lz);
  const long ox = 3 * n + (inx ? (long)j * 2 * lx + slab_pos(i, nx, lx) : 0);
  const long oz = 3 * n + 2 * SX + (inz ? (long)slab_pos(j, nz, lz) * nx + i : 0);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz);
  // Hy' feeds the E updates through dHy_dz (Ex node c, c+nx) and dHy_dx (Ez node c, c+1)
  const T lh = Lin[2 * n + c] + idz * (W[c] - (j + 1 < nz ? W[c + nx] : T(0))) +
               idx * (W[n + c] - (i + 1 < nx ? W[n + c + 1] : T(0)));
  T la = 0, lb = 0;
  if (i <= nx - 2 && j <= nz - 2) {  // Hy node, :249-266
    const bool ring = (i == 0 || i == nx - 2 || j == 0 || j == nz - 2);
    Lout[2 * n + c] = ring ? T(0) : lh;
    const T w = T(M.scal[0]) * lh;
    la = w;
    lb = w;
    if (inz) {
      const T lr = Lin[oz] + w;
      la = P[12 + PK][j] * w + P[12 + PA][j] * lr;
      Lout[oz] = P[12 + PB][j] * lr;
    }
    if (inx) {
      const T lr = Lin[ox] + w;
      lb = P[PK][i] * w + P[PA][i] * lr;
      Lout[ox] = P[PB][i] * lr;
    }
  } else {
    Lout[2 * n + c] = lh;
    if (inz) Lout[oz] = Lin[oz];
    if (inx) Lout[ox] = Lin[ox];
  }
  W[2 * n + c] = la;  // adjoint of dEx_dz at the Hy node
  W[3 * n + c] = lb;  // adjoint of dEz_dx at the Hy node
}

template <typename T>
__global__ void __launch_bounds__(NT)
k_em2_adj_g(const DModel* __restrict__ models, const DShot* __restrict__ shots,
            const DAdj* __restrict__ adj, int nshots, const int* __restrict__ tbase, int k) {
  const int s = blockIdx.z;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TZ + threadIdx.y;
  if (i >= nx || j >= nz) return;
  const int t = *tbase + k;
  const long n = (long)nx * nz, c = (long)j * nx + i;
  auto Lout = mp<T>(A.lam[t & 1]);
  const auto W = cp<T>(A.work);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz);
  Lout[c] = Lout[c] + idz * ((j >= 1 ? W[2 * n + c - nx] : T(0)) - W[2 * n + c]);
  Lout[n + c] = Lout[n + c] + idx * ((i >= 1 ? W[3 * n + c - 1] : T(0)) - W[3 * n + c]);
}

// ---------------------------------------------------------------------------
// Fused EM2 adjoint step: the transpose of k_em2_adj_{e,h,g} in one pass,
// the same arithmetic in the same order.  A wavefront owns AW columns
// (lanes 2..AW+1; lane l holds column xo - 2 + l) of eh rows of one shot and
// marches down them (eh rows: the plan picks the height that gives it
// enough tiles): at march row q it does E^T of row q, H^T of row q-1
// (which reads the E^T results of rows q-1 and q and of column i+1) and the
// E gather of row q-1 (the H^T results of rows q-2 and q-1 and of column
// i-1).  E^T at column i reads the new Hy of column i-1, so two halo lanes on
// the left and one on the right give exact results on the owned lanes; the
// halo row above the tile is recomputed, so tiles are independent (lambda
// ping-pongs).  Only owned cells write lambda_t and add to the gradient.
// ---------------------------------------------------------------------------
template <typename T, bool PML>
__device__ __forceinline__ void em2_adj_tile(const DModel* __restrict__ models,
                                             const DShot* __restrict__ shots,
                                             const DAdj* __restrict__ adj, const int s, const int xo,
                                             const int zo, const int eh, const int t, const int t0,
                                             const unsigned gmask, const int lane, const int tb) {
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const DAdj& A = adj[s];
  const int nx = M.nx, nz = M.nz;
  const int i = xo - 2 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);
  const bool own = incol && lane >= 2 && lane < 2 + AW;
  const int zend = min(zo + eh, nz);
  // 32-bit element offsets within one state / gradient array (checked at setup)
  const uint32_t n = (uint32_t)nx * nz;
  const int lx = M.lx, lz = M.lz;
  const uint32_t SX = (uint32_t)nz * 2 * lx, SZ = (uint32_t)2 * lz * nx;
  const A32<const T> Lin(A.lam[(t + 1) & 1]);
  const A32<T> Lout(A.lam[t & 1]);
  const A32<const T> u1(static_cast<const T*>(A.trail[tb]) + (long)(t + 1 - t0) * A.slot);
  const A32<const T> u0(static_cast<const T*>(A.trail[tb]) + (long)(t - t0) * A.slot);
  const A32<typename std::remove_pointer<decltype(GAcc<T>::of(A))>::type> G(GAcc<T>::of(A));
  const Prof<T> P(&M.prof[0][0]);
  const bool inx = PML && incol && in_slab(i, nx, lx);
  const uint32_t xs = inx ? slab_pos(i, nx, lx) : 0;
  const T ikx = P[PK][ic], ax = P[PA][ic], bx = P[PB][ic];
  const T ikxh = P[PKH][ic], axh = P[PAH][ic], bxh = P[PBH][ic];
  const auto ikz = P[12 + PK], az = P[12 + PA], bz = P[12 + PB];
  const auto ikzh = P[12 + PKH], azh = P[12 + PAH], bzh = P[12 + PBH];
  const A32<const T> CAX(M.coef[0]), CBX(M.coef[1]), CAZ(M.coef[2]), CBZ(M.coef[3]);
  const T idx = T(1.0 / M.dx), idz = T(1.0 / M.dz), dbh = T(M.scal[0]);
  auto row = [&](int r) { return (uint32_t)(r < 0 ? 0 : (r >= nz ? nz - 1 : r)) * nx + ic; };
  auto clampr = [&](int r) { return r < 0 ? 0 : (r >= nz ? nz - 1 : r); };

  // carried from the previous march row: E^T of row q-1 (adjoints of dHy_dz
  // at the Ex node, of dHy_dx at the Ez node, and the partial lambda_E), the
  // H^T result la of row q-2, and u_{t+1}'s Hy of row q-1
  T wx_p = 0, wz_p = 0, ex_p = 0, ez_p = 0, la_pp = 0;
  T hy1_p = (incol && zo - 2 >= 0) ? u1[2 * n + row(zo - 2)] : T(0);
  for (int q = zo - 1; q <= zend; ++q) {
    // ---- E^T of row q (k_em2_adj_e) ----
    const bool vq = incol && q >= 0 && q < nz;
    const int qs = clampr(q);
    const uint32_t c = row(q);
    const T hy1 = vq ? u1[2 * n + c] : T(0);
    const T hy1_w = lane_prev(hy1);  // u_{t+1} Hy(i-1, q)
    const bool ownq = own && q >= zo && q < zend;
    T wx = 0, wz = 0, exq = 0, ezq = 0;
    if (vq) {
      const bool inz = PML && in_slab(qs, nz, lz);
      const uint32_t ox = 3 * n + (inx ? (uint32_t)qs * 2 * lx + xs : 0);
      const uint32_t oz = 3 * n + 2 * SX + (inz ? (uint32_t)slab_pos(qs, nz, lz) * nx + ic : 0);
      // Dirichlet ring of Ex and Ez (and cells outside their Fortran extent)
      const bool edx = (i == 0 || i >= nx - 2 || q == 0 || q == nz - 1);
      const bool edz = (i == 0 || i == nx - 1 || q == 0 || q >= nz - 2);
      const T lex = edx ? T(0) : Lin[c], lez = edz ? T(0) : Lin[n + c];
      if (ownq && A.gsrc && i == S.isrc && q == S.ksrc) {
        A.gsrc[3 * (long)t + 0] = (double)lex;
        A.gsrc[3 * (long)t + 2] = (double)lez;
      }
      if (q >= 1 && q <= nz - 2 && i <= nx - 2) {  // Ex node, :273-283
        const T ld = CBX[c] * lex;
        const T d = (hy1 - hy1_p) * idz;
        T D = d;
        wx = ld;
        if (inz) {
          const T lr = Lin[oz + SZ] + ld;
          wx = ikzh[qs] * ld + azh[qs] * lr;
          if (ownq) Lout[oz + SZ] = bzh[qs] * lr;
          D = ikzh[qs] * d + u1[oz + SZ];
        }
        if (ownq) {
          if (gmask & 1u) gadd(&G[c], (double)lex * (double)u0[c]);
          if (gmask & 2u) gadd(&G[n + c], (double)lex * (double)D);
        }
        exq = CAX[c] * lex;
      } else {
        if (ownq && inz) Lout[oz + SZ] = Lin[oz + SZ];
        exq = lex;
      }
      if (q <= nz - 2 && i >= 1 && i <= nx - 2) {  // Ez node, :285-295
        const T le = CBZ[c] * lez;
        const T e = (hy1 - hy1_w) * idx;
        T E = e;
        wz = le;
        if (inx) {
          const T lr = Lin[ox + SX] + le;
          wz = ikxh * le + axh * lr;
          if (ownq) Lout[ox + SX] = bxh * lr;
          E = ikxh * e + u1[ox + SX];
        }
        if (ownq) {
          if (gmask & 4u) gadd(&G[2 * n + c], (double)lez * (double)u0[n + c]);
          if (gmask & 8u) gadd(&G[3 * n + c], (double)lez * (double)E);
        }
        ezq = CAZ[c] * lez;
      } else {
        if (ownq && inx) Lout[ox + SX] = Lin[ox + SX];
        ezq = lez;
      }
    }

    // ---- H^T of row r = q-1 (k_em2_adj_h) ----
    const int r = q - 1;
    const T wz_e = lane_next(wz_p);  // E^T adjoint of dHy_dx at the Ez node (i+1, r)
    T la = 0, lb = 0;
    const bool vr = incol && r >= 0 && r < nz && q >= zo;
    if (vr) {
      const bool ownr = own && r >= zo && r < zend;
      const uint32_t cr = row(r);
      const bool inz = PML && in_slab(r, nz, lz);
      const uint32_t ox = 3 * n + (inx ? (uint32_t)r * 2 * lx + xs : 0);
      const uint32_t oz = 3 * n + 2 * SX + (inz ? (uint32_t)slab_pos(r, nz, lz) * nx + ic : 0);
      // Hy' feeds the E updates through dHy_dz (Ex node c, c+nx) and dHy_dx (Ez node c, c+1)
      const T lh = Lin[2 * n + cr] + idz * (wx_p - (r + 1 < nz ? wx : T(0))) +
                   idx * (wz_p - (i + 1 < nx ? wz_e : T(0)));
      if (i <= nx - 2 && r <= nz - 2) {  // Hy node, :249-266
        const bool ring = (i == 0 || i == nx - 2 || r == 0 || r == nz - 2);
        if (ownr) Lout[2 * n + cr] = ring ? T(0) : lh;
        const T w = dbh * lh;
        la = w;
        lb = w;
        if (inz) {
          const T lr = Lin[oz] + w;
          la = ikz[r] * w + az[r] * lr;
          if (ownr) Lout[oz] = bz[r] * lr;
        }
        if (inx) {
          const T lr = Lin[ox] + w;
          lb = ikx * w + ax * lr;
          if (ownr) Lout[ox] = bx * lr;
        }
      } else if (ownr) {
        Lout[2 * n + cr] = lh;
        if (inz) Lout[oz] = Lin[oz];
        if (inx) Lout[ox] = Lin[ox];
      }
    }

    // ---- E gather of row r (k_em2_adj_g) ----
    const T lb_w = lane_prev(lb);  // H^T adjoint of dEz_dx at the Hy node (i-1, r)
    if (own && r >= zo && r < zend) {
      const uint32_t cr = row(r);
      Lout[cr] = ex_p + idz * ((r >= 1 ? la_pp : T(0)) - la);
      Lout[n + cr] = ez_p + idx * ((i >= 1 ? lb_w : T(0)) - lb);
    }

    // ---- shift the windows ----
    wx_p = wx; wz_p = wz; ex_p = exq; ez_p = ezq;
    la_pp = la;
    hy1_p = hy1;
  }
}

template <typename T>
__global__ void __launch_bounds__(64 * AWB)
k_em2_adj(const DModel* __restrict__ models, const DShot* __restrict__ shots,
          const DAdj* __restrict__ adj, const int4* __restrict__ work, int nitems,
          const int* __restrict__ tbase, int k, unsigned gmask, const int* __restrict__ tsel) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t0 = __builtin_amdgcn_readfirstlane(*tbase);
  const int tb = __builtin_amdgcn_readfirstlane(*tsel);
  const int t = t0 + k;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const int4 wv = work[w];
    const int s = __builtin_amdgcn_readfirstlane(wv.x);
    const int xo = __builtin_amdgcn_readfirstlane(wv.y) + wave * AW;
    const int zo = __builtin_amdgcn_readfirstlane(wv.z);
    const int eh = __builtin_amdgcn_readfirstlane(wv.w);  // tile height
    if (s < 0) continue;
    const DModel& M = models[shots[s].model];
    if (xo >= M.nx) continue;
    // a wave whose columns and rows (halo included) avoid every C-PML slab
    // runs the body without memory variables
    const bool interior = xo - 2 >= M.lx && xo + 61 < M.nx - M.lx && zo - 1 >= M.lz &&
                          zo + eh < M.nz - M.lz;
    if (interior)
      em2_adj_tile<T, false>(models, shots, adj, s, xo, zo, eh, t, t0, gmask, threadIdx.x & 63, tb);
    else
      em2_adj_tile<T, true>(models, shots, adj, s, xo, zo, eh, t, t0, gmask, threadIdx.x & 63, tb);
  }
}

// ===========================================================================
// EM25: electromag25, legacy/fortran/electromagfdtd.f95:619-1092
// psi slabs: x [nz][ny][2lx]: dEz_dx dEy_dx dHz_dx dHy_dx
//            y [nz][2ly][nx]: dEz_dy dEx_dy dHz_dy dHx_dy
//            z [2lz][ny][nx]: dEy_dz dEx_dz dHy_dz dHx_dz
// H boundary zeroing (:1059-1078) is applied by the next H update.
// Two launches per step (H, then E) in the layout of the 2.5-D elastic
// kernels: a wavefront takes one plane, 64 columns (62 owned; x neighbours by
// lane shuffles) and a chunk of rows, walking y with the y neighbours in
// registers; the z neighbours are loads of the adjacent plane.
// ===========================================================================
struct Slab3e {
  long SX, SY, SZ, bx, by, bz;
  __device__ Slab3e(int nx, int ny, int nz, int lx, int ly, int lz) {
    SX = (long)nz * ny * 2 * lx;
    SY = (long)nz * 2 * ly * nx;
    SZ = (long)2 * lz * ny * nx;
    bx = 0;
    by = 4 * SX;
    bz = 4 * SX + 4 * SY;
  }
};

template <typename T, bool PML>
__device__ __forceinline__ void em25_h_tile(const DModel* __restrict__ models,
                                           const DShot* __restrict__ shots, const int4 wv) {
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const int nx = M.nx, ny = M.ny, nz = M.nz;
  const int k = __builtin_amdgcn_readfirstlane(wv.z);
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(wv.y) - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);
  // H lives on i <= nx-2, j <= ny-2, k <= nz-2 (the loop ranges of :901-962)
  const bool own = incol && lane >= 1 && lane <= EW25 && i <= nx - 2 && k <= nz - 2;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w));
  const int j1 = min(ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w)), ny - 1);
  const long nxy = (long)nx * ny;
  const T idx = T(1.0 / M.dx), idy = T(1.0 / M.dy), idz = T(1.0 / M.dz), dbH = T(M.scal[0]);
  const int lx = M.lx, ly = M.ly, lz = M.lz;
  const bool inx = PML && incol && in_slab(i, nx, lx), inz = PML && in_slab(k, nz, lz);
  const Prof<T> P(&M.prof[0][0]);
  const T bx_i = P[PB][ic], ax_i = P[PA][ic], ikx_i = P[PK][ic];
  const T bz_k = P[12 + PB][k], az_k = P[12 + PA][k], ikz_k = P[12 + PK][k];
  const T bzh_k = P[12 + PBH][k], azh_k = P[12 + PAH][k], ikzh_k = P[12 + PKH][k];
  gcptr<T> by = P[6 + PB], ay = P[6 + PA], iky = P[6 + PK];
  gcptr<T> byh = P[6 + PBH], ayh = P[6 + PAH], ikyh = P[6 + PKH];
  const int ku = k + 1 < nz ? k + 1 : k;
  const A32<const T> EX = cplane<T>(S.f[0], nxy, k), EY = cplane<T>(S.f[1], nxy, k);
  const A32<const T> EZ = cplane<T>(S.f[2], nxy, k);
  const A32<const T> EXU = cplane<T>(S.f[0], nxy, ku), EYU = cplane<T>(S.f[1], nxy, ku);
  const A32<T> HX = plane<T>(S.f[3], nxy, k), HY = plane<T>(S.f[4], nxy, k), HZ = plane<T>(S.f[5], nxy, k);
  auto psi = mp<T>(S.psi);
  const Slab3e L(nx, ny, nz, lx, ly, lz);
  const int xs = inx ? slab_pos(i, nx, lx) : 0, zs = inz ? slab_pos(k, nz, lz) : 0;
  if (j0 >= j1) return;
  // y window: Ex, Ez of row j
  uint32_t c = (uint32_t)j0 * nx + ic;
  T exc = EX[c], ezc = EZ[c];
#pragma unroll 1
  for (int j = j0; j < j1; ++j, c += nx) {
    const T exn = EX[c + nx], ezn = EZ[c + nx];  // row j+1 (j <= ny-2)
    const T eyc = EY[c];
    const T eyu = EYU[c], exu = EXU[c];          // plane k+1
    const T eze = lane_next(ezc), eye = lane_next(eyc);
    const T hx0 = HX[c], hy0 = HY[c], hz0 = HZ[c];
    if (own) {
      const bool iny = PML && in_slab(j, ny, ly);
      const long ox = L.bx + ((long)k * ny + j) * 2 * lx + xs;
      const long oy = L.by + ((long)k * 2 * ly + (iny ? slab_pos(j, ny, ly) : 0)) * nx + ic;
      const long oz = L.bz + ((long)zs * ny + j) * nx + ic;
      const bool lo = (i == 0 || j == 0 || k == 0);
      {  // Hx, :901-920
        T a = (ezc - ezn) * idy;  // dEz_dy
        if (iny) { T m = byh[j] * psi[oy] + ayh[j] * a; a = a * ikyh[j] + m; psi[oy] = m; }
        T b = (eyu - eyc) * idz;  // dEy_dz
        if (inz) { T m = bzh_k * psi[oz] + azh_k * b; b = b * ikzh_k + m; psi[oz] = m; }
        const bool edge = lo || i == nx - 1 || j == ny - 2 || k == nz - 2;
        HX[c] = (edge ? T(0) : hx0) + dbH * (b + a);
      }
      {  // Hy, :922-942
        T a = (exc - exu) * idz;  // dEx_dz
        if (inz) { T m = bz_k * psi[oz + L.SZ] + az_k * a; a = a * ikz_k + m; psi[oz + L.SZ] = m; }
        T b = (eze - ezc) * idx;  // dEz_dx
        if (inx) { T m = bx_i * psi[ox] + ax_i * b; b = b * ikx_i + m; psi[ox] = m; }
        const bool edge = lo || i == nx - 2 || j == ny - 1 || k == nz - 2;
        HY[c] = (edge ? T(0) : hy0) + dbH * (b + a);
      }
      if (k >= 1) {  // Hz, :944-962
        T a = (exn - exc) * idy;  // dEx_dy
        if (iny) { T m = by[j] * psi[oy + L.SY] + ay[j] * a; a = a * iky[j] + m; psi[oy + L.SY] = m; }
        T b = (eyc - eye) * idx;  // dEy_dx
        if (inx) { T m = bx_i * psi[ox + L.SX] + ax_i * b; b = b * ikx_i + m; psi[ox + L.SX] = m; }
        const bool edge = lo || i == nx - 2 || j == ny - 2 || k == nz - 1;
        HZ[c] = (edge ? T(0) : hz0) + dbH * (b + a);
      }
    }
    exc = exn;
    ezc = ezn;
  }
}

template <typename T, bool PML>
__device__ __forceinline__ void em25_e_tile(const DModel* __restrict__ models,
                                           const DShot* __restrict__ shots, const int4 wv,
                                           const int it) {
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  const DShot& S = shots[s];
  const DModel& M = models[S.model];
  const int nx = M.nx, ny = M.ny, nz = M.nz;
  const int k = __builtin_amdgcn_readfirstlane(wv.z);
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(wv.y) - 1 + lane;
  const bool incol = (i >= 0 && i < nx);
  const int ic = incol ? i : (i < 0 ? 0 : nx - 1);
  const bool own = incol && lane >= 1 && lane <= EW25;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w));
  const int j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  const long nxy = (long)nx * ny;
  const uint32_t p = (uint32_t)k * nx + ic;
  const T idx = T(1.0 / M.dx), idy = T(1.0 / M.dy), idz = T(1.0 / M.dz);
  const int lx = M.lx, ly = M.ly, lz = M.lz;
  const bool inx = PML && incol && in_slab(i, nx, lx), inz = PML && in_slab(k, nz, lz);
  const Prof<T> P(&M.prof[0][0]);
  const T bxh_i = P[PBH][ic], axh_i = P[PAH][ic], ikxh_i = P[PKH][ic];
  const T bz_k = P[12 + PB][k], az_k = P[12 + PA][k], ikz_k = P[12 + PK][k];
  const T bzh_k = P[12 + PBH][k], azh_k = P[12 + PAH][k], ikzh_k = P[12 + PKH][k];
  gcptr<T> byh = P[6 + PBH], ayh = P[6 + PAH], ikyh = P[6 + PKH];
  // loop ranges of :968-1029
  const bool ux = own && k >= 1 && k <= nz - 2 && i <= nx - 2;
  const bool uy = own && k >= 1 && k <= nz - 2 && i >= 1 && i <= nx - 2;
  const bool uz = own && k <= nz - 2 && i >= 1 && i <= nx - 2;
  const T cax = ux ? A32<const T>(M.coef[0])[p] : T(0), cbx = ux ? A32<const T>(M.coef[1])[p] : T(0);
  const T cay = uy ? A32<const T>(M.coef[2])[p] : T(0), cby = uy ? A32<const T>(M.coef[3])[p] : T(0);
  const T caz = uz ? A32<const T>(M.coef[4])[p] : T(0), cbz = uz ? A32<const T>(M.coef[5])[p] : T(0);
  const int kd = k >= 1 ? k - 1 : 0;
  const A32<const T> HX = cplane<T>(S.f[3], nxy, k), HY = cplane<T>(S.f[4], nxy, k);
  const A32<const T> HZ = cplane<T>(S.f[5], nxy, k);
  const A32<const T> HXD = cplane<T>(S.f[3], nxy, kd), HYD = cplane<T>(S.f[4], nxy, kd);
  const A32<T> EX = plane<T>(S.f[0], nxy, k), EY = plane<T>(S.f[1], nxy, k), EZ = plane<T>(S.f[2], nxy, k);
  auto psi = mp<T>(S.psi);
  const Slab3e L(nx, ny, nz, lx, ly, lz);
  const int xs = inx ? slab_pos(i, nx, lx) : 0, zs = inz ? slab_pos(k, nz, lz) : 0;
  const bool srccol = (i == S.isrc && k == S.ksrc);
  // y window: Hx, Hz of row j-1
  uint32_t c = (uint32_t)j0 * nx + ic;
  T hxm = 0, hzm = 0;
  if (j0 >= 1) { hxm = HX[c - nx]; hzm = HZ[c - nx]; }
#pragma unroll 1
  for (int j = j0; j < j1; ++j, c += nx) {
    const T hxc = HX[c], hyc = HY[c], hzc = HZ[c];
    const T hyd = HYD[c], hxd = HXD[c];          // plane k-1
    const T hzw = lane_prev(hzc), hyw = lane_prev(hyc);
    T ex = EX[c], ey = EY[c], ez = EZ[c];
    if (own) {
      const bool iny = PML && in_slab(j, ny, ly);
      const long ox = L.bx + ((long)k * ny + j) * 2 * lx + xs;
      const long oy = L.by + ((long)k * 2 * ly + (iny ? slab_pos(j, ny, ly) : 0)) * nx + ic;
      const long oz = L.bz + ((long)zs * ny + j) * nx + ic;
      if (ux && j >= 1 && j <= ny - 2) {  // Ex, :968-985
        T a = (hzc - hzm) * idy;  // dHz_dy
        if (iny) { T m = byh[j] * psi[oy + 2 * L.SY] + ayh[j] * a; a = a * ikyh[j] + m; psi[oy + 2 * L.SY] = m; }
        T b = (hyd - hyc) * idz;  // dHy_dz
        if (inz) { T m = bz_k * psi[oz + 2 * L.SZ] + az_k * b; b = b * ikz_k + m; psi[oz + 2 * L.SZ] = m; }
        ex = cax * ex + cbx * (a + b);
      }
      if (uy && j <= ny - 2) {  // Ey, :988-1006
        T a = (hzw - hzc) * idx;  // dHz_dx
        if (inx) { T m = bxh_i * psi[ox + 2 * L.SX] + axh_i * a; a = a * ikxh_i + m; psi[ox + 2 * L.SX] = m; }
        T b = (hxc - hxd) * idz;  // dHx_dz
        if (inz) { T m = bzh_k * psi[oz + 3 * L.SZ] + azh_k * b; b = b * ikzh_k + m; psi[oz + 3 * L.SZ] = m; }
        ey = cay * ey + cby * (a + b);
      }
      if (uz && j >= 1 && j <= ny - 2) {  // Ez, :1011-1029
        T a = (hxm - hxc) * idy;  // dHx_dy
        if (iny) { T m = byh[j] * psi[oy + 3 * L.SY] + ayh[j] * a; a = a * ikyh[j] + m; psi[oy + 3 * L.SY] = m; }
        T b = (hyc - hyw) * idx;  // dHy_dx
        if (inx) { T m = bxh_i * psi[ox + 3 * L.SX] + axh_i * b; b = b * ikxh_i + m; psi[ox + 3 * L.SX] = m; }
        ez = caz * ez + cbz * (a + b);
      }
      if (srccol && j == S.jsrc) {  // source, :1033-1035
        if (S.src[0]) ex += cp<T>(S.src[0])[it];
        if (S.src[1]) ey += cp<T>(S.src[1])[it];
        if (S.src[2]) ez += cp<T>(S.src[2])[it];
      }
      // Dirichlet faces, :1038-1057 (and cells outside each staggered field's extent)
      const bool lo = (i == 0 || j == 0 || k == 0);
      if (lo || i >= nx - 2 || j == ny - 1 || k == nz - 1) ex = 0;
      if (lo || i == nx - 1 || j >= ny - 2 || k == nz - 1) ey = 0;
      if (lo || i == nx - 1 || j == ny - 1 || k >= nz - 2) ez = 0;
      check_blowup((double)ex * ex + (double)ey * ey + (double)ez * ez, S.blowup, it);
      EX[c] = ex;
      EY[c] = ey;
      EZ[c] = ez;
    }
    hxm = hxc;
    hzm = hzc;
  }
}

// A wavefront whose columns, plane and rows (with the neighbours it reads)
// lie outside every C-PML slab runs the body without memory variables.
__device__ __forceinline__ bool em25_interior(const DModel& M, int xo, int k, int j0, int j1) {
  return xo - 1 >= M.lx && xo + 62 < M.nx - M.lx && k >= M.lz && k < M.nz - M.lz &&
         j0 >= M.ly && j1 <= M.ny - M.ly;
}

template <typename T>
__global__ void __launch_bounds__(64)
k_em25_h(const DModel* __restrict__ models, const DShot* __restrict__ shots,
         const int4* __restrict__ work) {
  const int4 wv = work[blockIdx.x];
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  if (s < 0) return;
  const DModel& M = models[shots[s].model];
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  if (em25_interior(M, __builtin_amdgcn_readfirstlane(wv.y), __builtin_amdgcn_readfirstlane(wv.z), j0, j1))
    em25_h_tile<T, false>(models, shots, wv);
  else
    em25_h_tile<T, true>(models, shots, wv);
}

template <typename T>
__global__ void __launch_bounds__(64)
k_em25_e(const DModel* __restrict__ models, const DShot* __restrict__ shots,
         const int4* __restrict__ work, const int* __restrict__ base, int rel) {
  const int4 wv = work[blockIdx.x];
  const int s = __builtin_amdgcn_readfirstlane(wv.x);
  if (s < 0) return;
  const DModel& M = models[shots[s].model];
  const int it = *base + rel;
  const int j0 = ychunk_j0(__builtin_amdgcn_readfirstlane(wv.w)), j1 = ychunk_j1(__builtin_amdgcn_readfirstlane(wv.w));
  if (em25_interior(M, __builtin_amdgcn_readfirstlane(wv.y), __builtin_amdgcn_readfirstlane(wv.z), j0, j1))
    em25_e_tile<T, false>(models, shots, wv, it);
  else
    em25_e_tile<T, true>(models, shots, wv, it);
}

// ===========================================================================
// Receiver gather (replaces arraybuild.py:199-295 reading snapshot files):
// trace[it][comp][r] = field_comp[rec_r] after the step.
// ===========================================================================
template <typename T>
__global__ void k_gather(const DModel* __restrict__ models, const DShot* __restrict__ shots,
                         int nshots, int ncomp, const int* __restrict__ base, int rel) {
  const int s = blockIdx.y;
  if (s >= nshots) return;
  const DShot& S = shots[s];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= S.nrec || S.trace == nullptr) return;
  const DModel& M = models[S.model];
  const int it = *base + rel;
  const auto q = cp<int>(S.rec) + 3 * r;
  const long c = ((long)q[2] * M.ny + q[1]) * M.nx + q[0];
  auto tr = mp<T>(S.trace) + ((long)it * ncomp) * S.nrec + r;
  for (int f = 0; f < ncomp; ++f) tr[(long)f * S.nrec] = cp<T>(S.f[f])[c];
}

// The H zeroing that closes each legacy 2.5-D EM step (electromagfdtd.f95
// :1059-1078) is folded into the next H update; at the end of a run it is
// applied here (one wavefront per work item of the EM25 kernels) so the
// returned state equals the legacy state.
template <typename T>
__global__ void __launch_bounds__(64)
k_em25_edges(const DModel* __restrict__ models, const DShot* __restrict__ shots,
             const int4* __restrict__ work) {
  const int4 wv = work[blockIdx.x];
  if (wv.x < 0) return;
  const DShot& S = shots[wv.x];
  const DModel& M = models[S.model];
  const int nx = M.nx, ny = M.ny, nz = M.nz;
  const int lane = threadIdx.x & 63;
  const int i = wv.y - 1 + lane, k = wv.z;
  if (lane < 1 || lane > EW25 || i >= nx || k >= nz) return;
  auto Hx = mp<T>(S.f[3]);
  auto Hy = mp<T>(S.f[4]);
  auto Hz = mp<T>(S.f[5]);
  for (int j = ychunk_j0(wv.w); j < ychunk_j1(wv.w); ++j) {
    const long c = ((long)k * ny + j) * nx + i;
    const bool lo = (i == 0 || j == 0 || k == 0);
    if (lo || i == nx - 1 || j == ny - 2 || k == nz - 2) Hx[c] = T(0);
    if (lo || i == nx - 2 || j == ny - 1 || k == nz - 2) Hy[c] = T(0);
    if (lo || i == nx - 2 || j == ny - 2 || k == nz - 1) Hz[c] = T(0);
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
__global__ void k_set_int2(int* p, int v, int* p2, int v2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *p = v;
    *p2 = v2;
  }
}
__global__ void k_add_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = *p + v;
}
__global__ void k_fill_int(int* p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// ===========================================================================
// HBM probe: a copy of known traffic (4 n bytes read, 4 n written) whose
// loads and stores are one 4-byte element per lane, consecutive lanes on
// consecutive addresses -- the access width of the stencil kernels' loads.
// Measurement only: it calibrates the byte counters the stencils' traffic
// is read from.
// ===========================================================================
__global__ void __launch_bounds__(256) k_hbm_probe(const float* __restrict__ src,
                                                   float* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// ===========================================================================
// Plan
// ===========================================================================
struct Plan {
  int kind = 0, prec = 0, nmodels = 0, nshots = 0, nt = 0, ncomp = 2;
  bool split = false;  // SDT_SPLIT_CPML
  int maxrec = 0;
  std::vector<sdt_model> hmodels;
  std::vector<sdt_shot> hshots;
  DModel* dmodels = nullptr;
  DShot* dshots = nullptr;
  int4* dwork = nullptr;
  int nwork = 0;
  int el2_grid = 0;  // persistent grid of the ELASTIC2 step kernel
  int el2_wb = SDT_EL2_WAVES_PER_BLOCK;  // tile: wavefronts per block
  int el2_eh = SDT_EL2_ROWS;             // tile: rows
  int el2_order = 0;  // work-list order: 0 tiles dealt to the XCDs in turn, 1 column strips per XCD
  int spl = 1;                 // steps per launch of the ELASTIC2 step kernel (1 or 2)
  int4* dwork2 = nullptr;      // work list of the two-step kernel: interior blocks first
  int nwork2 = 0, nwork2i = 0;
  int el2x2_grid = 0, el2x2_grid_i = 0;
  int* dit0 = nullptr;         // first step of the current run of two-step launches
  int* drec = nullptr;         // receivers by row of every shot (DShot::rrow, rxid)
  bool graph_two = false;      // the cached graph holds two-step launches
  bool b_holds = false;        // buffer B holds the current state too (ELASTIC2 / EM2)
  int* dflags = nullptr;  // per shot first unstable step (INT_MAX = none)
  int* dbase = nullptr;   // step index of rel = 0
  long* dnel = nullptr;   // per shot: field elements (ping-pong copy-back)
  long* dnpsi = nullptr;  // per shot: psi elements
  hipStream_t last = nullptr;
  hipStream_t cap = nullptr;  // capture stream
  int graph_block = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  DAdj* dadj = nullptr;          // adjoint descriptors (sdt_adjoint_setup)
  std::vector<DAdj> hadj;
  int4* dadjwork = nullptr;      // work list of the fused adjoint kernel
  int* dtbase = nullptr;         // first step of the trail (adjoint kernels)
  int* dtsel_trail = nullptr;    // trail buffer the recomputation writes
  int* dtsel_adj = nullptr;      // trail buffer the adjoint reads
  // captured segments: (kind 0 trail / 1 adjoint, key) -> graph; kept until
  // setup or destroy so a replay in flight is never freed
  struct SegGraph {
    int kind;
    long long key;
    hipGraph_t g;
    hipGraphExec_t e;
  };
  std::vector<SegGraph> seg_graphs;
  std::vector<float*> g32;       // FP32 gradient staging per distinct grad array
  DFlush* dflush = nullptr;
  int nflush = 0;
  int nadjwork = 0;
  unsigned gmask = 0xfffu;       // gradient slots to accumulate
  int adj_flags = 0;
  int maxnx = 0, maxnz = 0;
};

int kind_fields(int kind) {
  switch (kind) {
    case SDT_ELASTIC2: return 5;
    case SDT_ELASTIC25: return 9;
    case SDT_EM2: return 3;
    case SDT_EM25: return 6;
  }
  return 0;
}
int kind_ncoef(int kind) {
  switch (kind) {
    case SDT_ELASTIC2: return 12;
    case SDT_ELASTIC25: return 39;
    case SDT_EM2: return 4;
    case SDT_EM25: return 6;
  }
  return 0;
}
bool kind_3d(int kind) { return kind == SDT_ELASTIC25 || kind == SDT_EM25; }
// kinds whose single-pass step reads state A or B and writes the other
bool pingpong(int kind) { return kind == SDT_ELASTIC2 || kind == SDT_EM2; }

// kind may carry SDT_SPLIT_CPML (ELASTIC2 only: two more memory arrays per slab)
long long psi_elems(int kind, long long nx, long long ny, long long nz, long long lx,
                    long long ly, long long lz) {
  if (kind & SDT_SPLIT_CPML) {
    if ((kind & ~SDT_SPLIT_CPML) != SDT_ELASTIC2) return -1;
    return 6 * nz * 2 * lx + 6 * 2 * lz * nx;
  }
  switch (kind) {
    case SDT_ELASTIC2: return 4 * nz * 2 * lx + 4 * 2 * lz * nx;
    case SDT_EM2: return 2 * nz * 2 * lx + 2 * 2 * lz * nx;
    case SDT_ELASTIC25: return 6 * (nz * ny * 2 * lx + nz * 2 * ly * nx + 2 * lz * ny * nx);
    case SDT_EM25: return 4 * (nz * ny * 2 * lx + nz * 2 * ly * nx + 2 * lz * ny * nx);
  }
  return -1;
}

// One launch of the single-pass ELASTIC2 step with the plan's tile shape.
template <typename T, bool TRAIL, bool SPLIT>
void launch_el2(Plan* p, hipStream_t st, int rel, const int* tsel, const int* prun = nullptr) {
  const DAdj* adj = TRAIL ? p->dadj : nullptr;
  const dim3 g(p->el2_grid);
  if (!TRAIL && p->el2_eh == 1) {  // tiles of one row: all loads at once
    switch (p->el2_wb) {
      case 1:
        hipLaunchKernelGGL((k_el2_step<T, false, SPLIT, 1, true>), g, dim3(64), 0, st, p->dmodels,
                           p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, 1, prun);
        break;
      case 2:
        hipLaunchKernelGGL((k_el2_step<T, false, SPLIT, 2, true>), g, dim3(128), 0, st, p->dmodels,
                           p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, 1, prun);
        break;
      default:
        hipLaunchKernelGGL((k_el2_step<T, false, SPLIT, 4, true>), g, dim3(256), 0, st, p->dmodels,
                           p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, 1, prun);
        break;
    }
    return;
  }
  switch (p->el2_wb) {
    case 1:
      hipLaunchKernelGGL((k_el2_step<T, TRAIL, SPLIT, 1>), g, dim3(64), 0, st, p->dmodels,
                         p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
    case 2:
      hipLaunchKernelGGL((k_el2_step<T, TRAIL, SPLIT, 2>), g, dim3(128), 0, st, p->dmodels,
                         p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
    default:
      hipLaunchKernelGGL((k_el2_step<T, TRAIL, SPLIT, 4>), g, dim3(256), 0, st, p->dmodels,
                         p->dshots, p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
  }
}

// One two-step ELASTIC2 step pair (steps base+rel, base+rel+1): the
// interior blocks, then the blocks that touch a C-PML slab or the outer ring.
template <typename T, bool SPLIT, bool INT>
void launch_el2x2_list(Plan* p, hipStream_t st, int rel, const int4* work, int n, int grid) {
  if (n <= 0) return;
  const dim3 g(grid);
  switch (p->el2_wb) {
    case 1:
      hipLaunchKernelGGL((k_el2_step2<T, SPLIT, 1, INT>), g, dim3(64), 0, st, p->dmodels, p->dshots,
                         work, n, p->dbase, rel, p->dit0, p->el2_eh);
      break;
    case 2:
      hipLaunchKernelGGL((k_el2_step2<T, SPLIT, 2, INT>), g, dim3(128), 0, st, p->dmodels, p->dshots,
                         work, n, p->dbase, rel, p->dit0, p->el2_eh);
      break;
    default:
      hipLaunchKernelGGL((k_el2_step2<T, SPLIT, 4, INT>), g, dim3(256), 0, st, p->dmodels, p->dshots,
                         work, n, p->dbase, rel, p->dit0, p->el2_eh);
      break;
  }
}
template <typename T, bool SPLIT>
void launch_el2x2(Plan* p, hipStream_t st, int rel) {
  launch_el2x2_list<T, SPLIT, true>(p, st, rel, p->dwork2, p->nwork2i, p->el2x2_grid_i);
  launch_el2x2_list<T, SPLIT, false>(p, st, rel, p->dwork2 + p->nwork2i, p->nwork2 - p->nwork2i,
                                     p->el2x2_grid);
}

template <typename T, bool TRAIL>
void launch_em2(Plan* p, hipStream_t st, int rel, const int* tsel, const int* prun = nullptr) {
  const DAdj* adj = TRAIL ? p->dadj : nullptr;
  const dim3 g(p->el2_grid);
  switch (p->el2_wb) {
    case 1:
      hipLaunchKernelGGL((k_em2_step<T, TRAIL, 1>), g, dim3(64), 0, st, p->dmodels, p->dshots,
                         p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
    case 2:
      hipLaunchKernelGGL((k_em2_step<T, TRAIL, 2>), g, dim3(128), 0, st, p->dmodels, p->dshots,
                         p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
    default:
      hipLaunchKernelGGL((k_em2_step<T, TRAIL, 4>), g, dim3(256), 0, st, p->dmodels, p->dshots,
                         p->dwork, p->nwork, p->dbase, rel, adj, tsel, p->el2_eh, prun);
      break;
  }
}

// One two-step EM2 launch (steps base+rel, base+rel+1).
template <typename T>
void launch_em2x2(Plan* p, hipStream_t st, int rel) {
  const dim3 g(p->el2x2_grid);
  const int4* w = p->dwork2;
  const int n = p->nwork2;
  switch (p->el2_wb) {
    case 1:
      hipLaunchKernelGGL((k_em2_step2<T, 1>), g, dim3(64), 0, st, p->dmodels, p->dshots, w, n, p->dbase,
                         rel, p->dit0, p->el2_eh);
      break;
    case 2:
      hipLaunchKernelGGL((k_em2_step2<T, 2>), g, dim3(128), 0, st, p->dmodels, p->dshots, w, n, p->dbase,
                         rel, p->dit0, p->el2_eh);
      break;
    default:
      hipLaunchKernelGGL((k_em2_step2<T, 4>), g, dim3(256), 0, st, p->dmodels, p->dshots, w, n, p->dbase,
                         rel, p->dit0, p->el2_eh);
      break;
  }
}

// Work list of the ELASTIC2 step kernel for the plan's tile shape: one item
// (shot, x0, z0) per block of el2_wb adjacent x tiles; the items of one
// block-tile over all shots of its model share blockIdx % 8 (one XCD's L2).
// Whether every wavefront of the two-step kernel's block at (x0, z0) takes
// the interior body (the test of k_el2_step2 for each of its wavefronts).
bool el2x2_interior(const sdt_model& M, int x0, int z0, int wb, int eh) {
  const int ze = std::min(z0 + eh, M.nz);
  if (!(z0 - 4 >= M.lz + 1 && ze + 1 <= M.nz - M.lz - 2)) return false;
  for (int w = 0; w < wb; ++w) {
    const int xw = x0 + w * EW2;
    if (xw >= M.nx) break;
    if (!(xw - 2 >= M.lx + 1 && xw + 61 <= M.nx - M.lx - 2)) return false;
  }
  return true;
}

// Default rows per tile of the 2-D step kernels: the tallest of 32, 16, 8, 4
// rows that still gives the plan 1024 blocks or more (4 per CU), else 2 --
// a single shot on a small grid needs short tiles to fill the device.
int default_rows(const Plan* p) {
  const int bw = EW * p->el2_wb;
  for (int rows : {32, 16, 8, 4}) {
    long long blocks = 0;
    for (const sdt_shot& S : p->hshots) {
      const sdt_model& M = p->hmodels[S.model];
      blocks += (long long)((M.nx + bw - 1) / bw) * ((M.nz + rows - 1) / rows);
    }
    if (blocks >= 1024) return rows;
  }
  return 2;
}

// Row cuts of one model's tiles: rows of at most eh, and with zt < zb also
// cuts at zt and zb, so the tiles between (those the kernel can run with
// its interior body) start exactly where that first becomes possible.
std::vector<int> row_cuts(int nz, int eh, int zt, int zb) {
  std::vector<int> c = {0};
  auto run = [&](int a, int b) {
    for (int z = a + eh; z < b; z += eh) c.push_back(z);
    if (b > c.back()) c.push_back(b);
  };
  if (zt > 0 && zt < zb && zb < nz) {
    run(0, zt);
    run(zt, zb);
    run(zb, nz);
  } else {
    run(0, nz);
  }
  return c;
}

// which: 0 every block, 1 the interior blocks of the two-step kernel, 2 the
// others.  ew: owned columns per wavefront; two-step lists (ew = EW2, EW2E)
// cut their rows at the interior limits of their kernel and carry each
// tile's height in .w (0: the plan's rows); one-step lists (ew = EW) carry
// the shot's model index in .w.  Order (p->el2_order): 0 deals a model's
// tiles to the XCDs in turn (tile k of the list on XCD k % 8), every XCD
// covering the whole model; 1 lists them column by column and gives each
// XCD a contiguous run, so a tile and the one below it, which share halo
// rows, run on one XCD one after the other, and each XCD keeps to one part
// of the model from step to step.  Either way all shots of a tile run on
// one XCD, one after another.
std::vector<int4> el2_work(const Plan* p, int ew = EW, int which = 0) {
  std::vector<int4> work;
  const int bw = ew * p->el2_wb, eh = p->el2_eh;
  for (int m = 0; m < p->nmodels; ++m) {
    const sdt_model& M = p->hmodels[m];
    // first and last-plus-one interior rows of k_el2_step2 / k_em2_step2
    const int zt = ew == EW2 ? M.lz + 5 : ew == EW2E ? M.lz + 3 : 0;
    const int zb = ew == EW2 ? M.nz - M.lz - 3 : ew == EW2E ? M.nz - M.lz - 4 : 0;
    const std::vector<int> cuts = row_cuts(M.nz, eh, zt, zb);
    const int tx = (M.nx + bw - 1) / bw, tz = (int)cuts.size() - 1, ntile = tx * tz;
    std::vector<int> ms;
    for (int s = 0; s < p->nshots; ++s)
      if (p->hshots[s].model == m) ms.push_back(s);
    if (ms.empty()) continue;
    std::vector<int> tl;
    auto take = [&](int t) {
      const int z0 = cuts[t / tx], h = cuts[t / tx + 1] - z0;
      if (which == 0 || el2x2_interior(M, (t % tx) * bw, z0, p->el2_wb, h) == (which == 1))
        tl.push_back(t);
    };
    if (p->el2_order) {
      for (int x = 0; x < tx; ++x)
        for (int z = 0; z < tz; ++z) take(z * tx + x);
    } else {
      for (int t = 0; t < ntile; ++t) take(t);
    }
    const int nt = (int)tl.size();
    // tile of slot tg of XCD tt (-1: none)
    const int chunk = (nt + 7) / 8;
    auto tile_at = [&](int tg, int tt) {
      const int k = p->el2_order ? tt * chunk + tg : tg * 8 + tt;
      return k < nt ? tl[k] : -1;
    };
    for (int tg = 0; tg < chunk; ++tg)
      for (int s : ms)
        for (int tt = 0; tt < 8; ++tt) {
          const int t = tile_at(tg, tt);
          if (t < 0) {
            work.push_back(make_int4(-1, 0, 0, 0));
            continue;
          }
          const int z0 = cuts[t / tx], h = cuts[t / tx + 1] - z0;
          work.push_back(make_int4(s, (t % tx) * bw, z0, ew == EW ? m : h));
        }
  }
  return work;
}

// Work list of the ELASTIC25 kernels: one item (shot, x0, k0, j0 | j1 << 16)
// per block of KB25 planes k0.. x XW25 * EW25 columns x0.. x rows j0..j1-1
// (at most Y25, cut at the y-slab edges), x fastest.
// The items of each shot are cut into 8 contiguous runs, one per XCD (item w
// runs on XCD w % 8), so the blocks an XCD runs at a time cover whole rows
// of a few planes (long contiguous runs of addresses in each array).
std::vector<int4> el25_work(const Plan* p, int planes, int xstep) {
  std::vector<int4> all;
  for (int s = 0; s < p->nshots; ++s) {
    const sdt_model& M = p->hmodels[p->hshots[s].model];
    // y chunks of at most Y25 rows, cut at the y-slab edges
    std::vector<int> cuts = {0};
    const int ya = std::min(M.ly, M.ny), yb = std::max(ya, M.ny - M.ly);
    if (ya > 0) cuts.push_back(ya);
    for (int j = ya + Y25; j < yb; j += Y25) cuts.push_back(j);
    if (yb > cuts.back()) cuts.push_back(yb);
    if (M.ny > cuts.back()) cuts.push_back(M.ny);
    for (size_t c = 0; c + 1 < cuts.size(); ++c)
      for (int k0 = 0; k0 < M.nz; k0 += planes)
        for (int x0 = 0; x0 < M.nx; x0 += xstep)
          all.push_back(make_int4(s, x0, k0, ychunk_pack(cuts[c], cuts[c + 1])));
  }
  const size_t run = (all.size() + 7) / 8;
  std::vector<int4> work(run * 8, make_int4(-1, 0, 0, 0));
  for (size_t q = 0; q < all.size(); ++q) work[(q % run) * 8 + q / run] = all[q];
  return work;
}

// Launch kernel `which` (0: first field update, 1: second field update,
// 2: receiver gather) of one step.
template <typename T>
void launch_one(Plan* p, hipStream_t st, int rel, int which) {
  const dim3 grd(p->nwork);
  if (which == 2) {
    // the 2-D step kernels write their receiver samples themselves
    if (p->maxrec > 0 && kind_3d(p->kind)) {
      const dim3 gg((p->maxrec + 63) / 64, p->nshots);
      hipLaunchKernelGGL(k_gather<T>, gg, dim3(64), 0, st, p->dmodels, p->dshots, p->nshots,
                         p->ncomp, p->dbase, rel);
    }
    return;
  }
  switch (p->kind * 2 + which) {
    case SDT_ELASTIC2 * 2:
      if (p->split)
        launch_el2<T, false, true>(p, st, rel, nullptr);
      else
        launch_el2<T, false, false>(p, st, rel, nullptr);
      break;
    case SDT_ELASTIC2 * 2 + 1:
      break;  // the fused step kernel does both phases
    case SDT_EM2 * 2:
      launch_em2<T, false>(p, st, rel, nullptr);
      break;
    case SDT_EM2 * 2 + 1:
      break;  // the fused step kernel does both phases
    case SDT_ELASTIC25 * 2:
      hipLaunchKernelGGL(k_el25_stress<T>, grd, dim3(NT25), 0, st, p->dmodels, p->dshots, p->dwork,
                         p->dbase, rel);
      break;
    case SDT_ELASTIC25 * 2 + 1:
      hipLaunchKernelGGL(k_el25_vel<T>, grd, dim3(NT25), 0, st, p->dmodels, p->dshots, p->dwork,
                         p->dbase, rel);
      break;
    case SDT_EM25 * 2:
      hipLaunchKernelGGL(k_em25_h<T>, grd, dim3(64), 0, st, p->dmodels, p->dshots, p->dwork);
      break;
    case SDT_EM25 * 2 + 1:
      hipLaunchKernelGGL(k_em25_e<T>, grd, dim3(64), 0, st, p->dmodels, p->dshots, p->dwork, p->dbase,
                         rel);
      break;
  }
}

template <typename T>
void launch_copy_back(Plan* p, hipStream_t st, int to_b = 0) {
  hipLaunchKernelGGL(k_copy_back<T>, dim3(1024, p->nshots), dim3(256), 0, st, p->dshots,
                     kind_fields(p->kind), p->dnel, p->dnpsi, to_b);
}

template <typename T>
void launch_finalize(Plan* p, hipStream_t st) {
  if (p->kind == SDT_EM2)
    hipLaunchKernelGGL(k_em2_edges<T>, dim3(16, p->nshots), dim3(256), 0, st, p->dmodels, p->dshots);
  if (p->kind == SDT_EM25)
    hipLaunchKernelGGL(k_em25_edges<T>, dim3(p->nwork), dim3(64), 0, st, p->dmodels, p->dshots,
                       p->dwork);
}

template <typename T>
void launch_step(Plan* p, hipStream_t st, int rel) {
  for (int w = 0; w < 3; ++w) launch_one<T>(p, st, rel, w);
}

void launch_steps(Plan* p, hipStream_t st, int n) {
  for (int r = 0; r < n; ++r) {
    if (p->prec == SDT_FP64)
      launch_step<double>(p, st, r);
    else
      launch_step<float>(p, st, r);
  }
}

// Launch `which` of a run of two-step launches from the base step: pair
// r/2 (steps r, r+1) for even r, or the single step r that ends an odd
// run (one-step kernel, reading the buffer the pairs left the state in).
template <typename T>
void launch_two(Plan* p, hipStream_t st, int r, bool single) {
  const bool em = p->kind == SDT_EM2;
  if (!single) {
    if (em)
      launch_em2x2<T>(p, st, r);
    else if (p->split)
      launch_el2x2<T, true>(p, st, r);
    else
      launch_el2x2<T, false>(p, st, r);
    return;
  }
  if (em)
    launch_em2<T, false>(p, st, r, nullptr, p->dit0);
  else if (p->split)
    launch_el2<T, false, true>(p, st, r, nullptr, p->dit0);
  else
    launch_el2<T, false, false>(p, st, r, nullptr, p->dit0);
}

void launch_pairs(Plan* p, hipStream_t st, int n) {
  for (int r = 0; r < n; r += 2) {
    const bool single = r + 1 == n;
    if (p->prec == SDT_FP64)
      launch_two<double>(p, st, r, single);
    else
      launch_two<float>(p, st, r, single);
  }
}

// Whether a run of n steps in graph blocks of graph_block uses two-step
// launches (the blocks must hold whole pairs).
bool use_two(const Plan* p, int n, int graph_block) {
  return p->spl == 2 && n >= 2 && (graph_block <= 1 || n < graph_block || graph_block % 2 == 0);
}

int validate(int kind, bool split, int prec, const sdt_model* models, int nmodels,
             const sdt_shot* shots, int nshots, int nt) {
  const int kpsi = kind | (split ? SDT_SPLIT_CPML : 0);
  if (kind < 0 || kind > 3) return fail(SDT_EINVAL, "unknown kind");
  if (prec != SDT_FP64 && prec != SDT_FP32) return fail(SDT_EINVAL, "unknown precision");
  if (!models || nmodels <= 0 || !shots || nshots <= 0) return fail(SDT_EINVAL, "no models or shots");
  if (nt <= 0) return fail(SDT_EINVAL, "nt must be positive");
  const bool d3 = kind_3d(kind);
  for (int m = 0; m < nmodels; ++m) {
    const sdt_model& M = models[m];
    const std::string tag = "model " + std::to_string(m) + ": ";
    if (M.nx < 3 || M.nz < 3 || (d3 && M.ny < 3) || (!d3 && M.ny != 1))
      return fail(SDT_EINVAL, tag + "grid too small (nx, nz >= 3; ny >= 3 in 2.5-D, 1 in 2-D)");
    if ((long long)M.nx * M.ny * M.nz > (1LL << 40)) return fail(SDT_EINVAL, tag + "grid too large");
    if (M.lx < 0 || M.lz < 0 || M.lx > (M.nx + 1) / 2 || M.lz > (M.nz + 1) / 2 ||
        (d3 && (M.ly < 0 || M.ly > (M.ny + 1) / 2)))
      return fail(SDT_EINVAL, tag + "slab width out of range");
    if (!(M.dx > 0) || !(M.dz > 0) || (d3 && !(M.dy > 0)) || !(M.dt > 0))
      return fail(SDT_EINVAL, tag + "dx, dy, dz, dt must be positive");
    for (int q = 0; q < kind_ncoef(kind); ++q) {
      const bool optional = kind == SDT_ELASTIC2 && (q == 2 || q == 4 || q == 5 || q == 6 || q >= 10);
      if (!M.coef[q] && !optional)
        return fail(SDT_EINVAL, tag + "coefficient slot " + std::to_string(q) + " is null");
    }
    if (kind == SDT_ELASTIC2) {
      const bool m1 = M.coef[2] || M.coef[4] || M.coef[5] || M.coef[6];
      const bool m2 = M.coef[2] && M.coef[4] && M.coef[5] && M.coef[6];
      if (m1 != m2 || (!M.coef[10]) != (!M.coef[11]))
        return fail(SDT_EINVAL, tag + "ELASTIC2 slots 2,4,5,6 (and 10,11) must be all set or all null");
    }
    for (int a = 0; a < 3; ++a) {
      if (a == 1 && !d3) continue;
      for (int q = 0; q < 6; ++q)
        if (!M.prof[a][q]) return fail(SDT_EINVAL, tag + "C-PML profile missing");
    }
    if (kind == SDT_ELASTIC2 && !split && !M.aux)
      return fail(SDT_EINVAL, tag + "aux (b_x_half by row) missing");
    if (kind == SDT_ELASTIC2 || kind == SDT_EM2) {
      // the step kernel addresses each array with 32-bit byte offsets
      const long long es = prec == SDT_FP64 ? 8 : 4;
      const long long np = psi_elems(kpsi, M.nx, M.ny, M.nz, M.lx, M.ly, M.lz);
      if ((long long)M.nx * M.nz * es >= (1LL << 32) || np * es >= (1LL << 32))
        return fail(SDT_EINVAL, tag + "grid too large (every array must stay under 4 GiB)");
    }
    if (kind == SDT_ELASTIC25) {
      // the kernels address one z plane of a field with 32-bit byte offsets
      const long long es = prec == SDT_FP64 ? 8 : 4;
      if ((long long)M.nx * M.ny * es >= (1LL << 32) || (long long)M.nx * M.nz * es >= (1LL << 32))
        return fail(SDT_EINVAL, tag + "grid too large (an x-y plane must stay under 4 GiB)");
      if (M.ny > 65535) return fail(SDT_EINVAL, tag + "ny above 65535");
    }
  }
  for (int s = 0; s < nshots; ++s) {
    const sdt_shot& S = shots[s];
    const std::string tag = "shot " + std::to_string(s) + ": ";
    if (S.model < 0 || S.model >= nmodels) return fail(SDT_EINVAL, tag + "bad model index");
    const sdt_model& M = models[S.model];
    for (int f = 0; f < kind_fields(kind); ++f)
      if (!S.field[f]) return fail(SDT_EINVAL, tag + "field pointer null");
    if (!S.psi && psi_elems(kpsi, M.nx, M.ny, M.nz, M.lx, M.ly, M.lz) > 0)
      return fail(SDT_EINVAL, tag + "psi workspace null");
    if (pingpong(kind)) {
      for (int f = 0; f < kind_fields(kind); ++f)
        if (!S.field_b[f]) return fail(SDT_EINVAL, tag + "field_b pointer null (ELASTIC2 and EM2 ping-pong)");
      if (!S.psi_b && psi_elems(kpsi, M.nx, M.ny, M.nz, M.lx, M.ly, M.lz) > 0)
        return fail(SDT_EINVAL, tag + "psi_b workspace null");
    }
    if (S.isrc < 0 || S.isrc >= M.nx || S.ksrc < 0 || S.ksrc >= M.nz ||
        (d3 && (S.jsrc < 0 || S.jsrc >= M.ny)))
      return fail(SDT_EINVAL, tag + "source outside the padded grid");
    if (S.nrec < 0 || (S.nrec > 0 && (!S.rec || !S.trace)))
      return fail(SDT_EINVAL, tag + "receivers need rec and trace buffers");
  }
  return SDT_OK;
}

// State after the last step of a run: EM H edges zeroed; ping-pong kinds
// copied back into buffer A when the last step wrote buffer B (one-step runs
// ending before an odd step; final_b: whether a run of two-step launches
// ended in B, else -1).  A always holds the state afterwards; b_holds
// records whether B does too (it did when the last step wrote B).
void end_of_run(Plan* p, hipStream_t st, int next_step, int final_b = -1) {
  const bool b = pingpong(p->kind) && (final_b < 0 ? (next_step & 1) != 0 : final_b != 0);
  p->b_holds = b;
  if (p->prec == SDT_FP64) {
    if (b) launch_copy_back<double>(p, st);
    launch_finalize<double>(p, st);
  } else {
    if (b) launch_copy_back<float>(p, st);
    launch_finalize<float>(p, st);
  }
}

// Before a run: the one-step kernels read the state of step it from buffer
// it & 1, so a one-step run that starts at an odd step needs the state in B.
// Copy A into B unless the previous run left it there (a run of two-step
// launches always starts from A and needs nothing).
void start_of_run(Plan* p, hipStream_t st, int it0, bool two) {
  if (!pingpong(p->kind) || two || !(it0 & 1) || p->b_holds) return;
  if (p->prec == SDT_FP64)
    launch_copy_back<double>(p, st, 1);
  else
    launch_copy_back<float>(p, st, 1);
  p->b_holds = true;
}

long long state_elems(int kind, long long nx, long long ny, long long nz, long long lx,
                      long long ly, long long lz) {
  const long long np = psi_elems(kind, nx, ny, nz, lx, ly, lz);
  return np < 0 ? -1 : kind_fields(kind & ~SDT_SPLIT_CPML) * nx * ny * nz + np;
}

template <typename T>
void launch_trail_steps(Plan* p, hipStream_t st, int it0, int nsteps) {
  for (int r = 0; r < nsteps; ++r) {
    if (p->split)
      launch_el2<T, true, true>(p, st, r, p->dtsel_trail);
    else
      launch_el2<T, true, false>(p, st, r, p->dtsel_trail);
  }
}

// Adjoint steps t = tbase + k for k = khi-1 down to klo (tbase = the trail's
// first step, held on the device so a captured segment replays anywhere).
template <typename T, bool SPLIT>
void launch_adjoint_steps(Plan* p, hipStream_t st, int khi, int klo) {
  const dim3 blk(TX, TZ);
  const dim3 grd((p->maxnx + TX - 1) / TX, (p->maxnz + TZ - 1) / TZ, p->nshots);
  const int* tb = p->dtbase;
  for (int k = khi - 1; k >= klo; --k) {
    if (p->maxrec > 0)
      hipLaunchKernelGGL(k_el2_adj_inject<T>, dim3((p->maxrec + 63) / 64, p->nshots), dim3(64), 0,
                         st, p->dmodels, p->dshots, p->dadj, p->nshots, tb, k);
    if (!(p->adj_flags & SDT_ADJ_REFERENCE)) {
      hipLaunchKernelGGL((k_el2_adj<T, SPLIT>), dim3(p->nadjwork), dim3(64 * AWB), 0, st,
                         p->dmodels, p->dshots, p->dadj, p->dadjwork, p->nadjwork, tb, k, p->gmask,
                         p->dtsel_adj);
      continue;
    }
    hipLaunchKernelGGL((k_el2_adj_v<T, SPLIT>), grd, blk, 0, st, p->dmodels, p->dshots, p->dadj,
                       p->nshots, tb, k, p->dtsel_adj);
    hipLaunchKernelGGL((k_el2_adj_s<T, SPLIT>), grd, blk, 0, st, p->dmodels, p->dshots, p->dadj,
                       p->nshots, tb, k, p->dtsel_adj);
    hipLaunchKernelGGL(k_el2_adj_g<T>, grd, blk, 0, st, p->dmodels, p->dshots, p->dadj, p->nshots, tb, k);
  }
}

template <typename T>
void launch_trail_em2(Plan* p, hipStream_t st, int nsteps) {
  for (int r = 0; r < nsteps; ++r) launch_em2<T, true>(p, st, r, p->dtsel_trail);
}

void launch_trail(Plan* p, hipStream_t st, int nsteps) {
  if (p->kind == SDT_EM2) {
    if (p->prec == SDT_FP64)
      launch_trail_em2<double>(p, st, nsteps);
    else
      launch_trail_em2<float>(p, st, nsteps);
    return;
  }
  if (p->prec == SDT_FP64)
    launch_trail_steps<double>(p, st, 0, nsteps);
  else
    launch_trail_steps<float>(p, st, 0, nsteps);
}

template <typename T>
void launch_adjoint_em2(Plan* p, hipStream_t st, int khi, int klo) {
  const dim3 blk(TX, TZ);
  const dim3 grd((p->maxnx + TX - 1) / TX, (p->maxnz + TZ - 1) / TZ, p->nshots);
  const int* tb = p->dtbase;
  for (int k = khi - 1; k >= klo; --k) {
    if (p->maxrec > 0)
      hipLaunchKernelGGL(k_el2_adj_inject<T>, dim3((p->maxrec + 63) / 64, p->nshots), dim3(64), 0,
                         st, p->dmodels, p->dshots, p->dadj, p->nshots, tb, k);
    if (!(p->adj_flags & SDT_ADJ_REFERENCE)) {
      hipLaunchKernelGGL(k_em2_adj<T>, dim3(p->nadjwork), dim3(64 * AWB), 0, st, p->dmodels,
                         p->dshots, p->dadj, p->dadjwork, p->nadjwork, tb, k, p->gmask, p->dtsel_adj);
      continue;
    }
    hipLaunchKernelGGL(k_em2_adj_e<T>, grd, blk, 0, st, p->dmodels, p->dshots, p->dadj, p->nshots,
                       tb, k, p->dtsel_adj, p->gmask);
    hipLaunchKernelGGL(k_em2_adj_h<T>, grd, blk, 0, st, p->dmodels, p->dshots, p->dadj, p->nshots,
                       tb, k);
    hipLaunchKernelGGL(k_em2_adj_g<T>, grd, blk, 0, st, p->dmodels, p->dshots, p->dadj, p->nshots,
                       tb, k);
  }
}

void launch_adjoint(Plan* p, hipStream_t st, int khi, int klo) {
  if (p->kind == SDT_EM2) {
    if (p->prec == SDT_FP64)
      launch_adjoint_em2<double>(p, st, khi, klo);
    else
      launch_adjoint_em2<float>(p, st, khi, klo);
  } else if (p->prec == SDT_FP64)
    p->split ? launch_adjoint_steps<double, true>(p, st, khi, klo)
             : launch_adjoint_steps<double, false>(p, st, khi, klo);
  else
    p->split ? launch_adjoint_steps<float, true>(p, st, khi, klo)
             : launch_adjoint_steps<float, false>(p, st, khi, klo);
  if (p->nflush > 0)
    hipLaunchKernelGGL(k_grad_flush, dim3(256, p->nflush), dim3(256), 0, st, p->dflush, p->gmask);
}

void drop_adjoint_graphs(Plan* p) {
  for (auto& x : p->seg_graphs) {
    (void)hipGraphExecDestroy(x.e);
    (void)hipGraphDestroy(x.g);
  }
  p->seg_graphs.clear();
}

// Replay (capturing on first use) the segment graph (kind, key) on st.
// (Re)build the work list of the two-step ELASTIC2 kernel for the plan's
// tile shape.
// Work list of the two-step kernels for the plan's tile shape.  ELASTIC2:
// interior blocks first (their count in *ni), then the others; EM2: one list.
std::vector<int4> build_work2(const Plan* p, int* ni) {
  const bool em = p->kind == SDT_EM2;
  std::vector<int4> work = em ? std::vector<int4>() : el2_work(p, EW2, 1);
  *ni = (int)work.size();
  const std::vector<int4> edge = em ? el2_work(p, EW2E, 0) : el2_work(p, EW2, 2);
  work.insert(work.end(), edge.begin(), edge.end());
  if (work.empty()) work.push_back(make_int4(-1, 0, 0, 0));
  return work;
}

// Grid sizes of the two-step launches for a list of nw items, ni interior.
void set_work2_grids(Plan* p, int nw, int ni) {
  p->nwork2 = nw;
  p->nwork2i = ni;
  p->el2x2_grid_i = std::max(8, std::min(ni, SDT_EL2_GRID) / 8 * 8);
  p->el2x2_grid = std::max(8, std::min(p->nwork2 - ni, SDT_EL2_GRID) / 8 * 8);
}

int upload_work2(Plan* p) {
  int ni = 0;
  const std::vector<int4> work = build_work2(p, &ni);
  int4* dw = nullptr;
  HIP_TRY(hipMalloc(&dw, sizeof(int4) * work.size()));
  if (hipMemcpy(dw, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dw);
    return fail(SDT_EDEVICE, "hipMemcpy failed (work list)");
  }
  if (p->dwork2) (void)hipFree(p->dwork2);
  p->dwork2 = dw;
  set_work2_grids(p, (int)work.size(), ni);
  return SDT_OK;
}

template <typename F>
int replay_segment(Plan* p, int kind, long long key, hipStream_t st, F&& enqueue) {
  for (auto& x : p->seg_graphs)
    if (x.kind == kind && x.key == key) {
      HIP_TRY(hipGraphLaunch(x.e, st));
      return SDT_OK;
    }
  if (p->seg_graphs.size() >= 8) {  // many segment shapes: free the old ones once idle
    HIP_TRY(hipDeviceSynchronize());
    drop_adjoint_graphs(p);
  }
  Plan::SegGraph x{kind, key, nullptr, nullptr};
  HIP_TRY(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
  enqueue(p->cap);
  HIP_TRY(hipStreamEndCapture(p->cap, &x.g));
  HIP_TRY(hipGraphInstantiate(&x.e, x.g, nullptr, nullptr, 0));
  p->seg_graphs.push_back(x);
  HIP_TRY(hipGraphLaunch(x.e, st));
  return SDT_OK;
}

// What a plan's device arrays hold, built on the host: model and shot
// descriptors, work lists, the receivers-by-row index and the per-shot
// sizes.  The descriptors' pointers into the plan's own arrays (blowup,
// rrow, rxid) are set by stage_bind once those arrays have addresses.
struct Stage {
  std::vector<DModel> dm;
  std::vector<DShot> ds;
  std::vector<int4> work;       // one-step ELASTIC2 / EM2 list, or the 2.5-D list
  std::vector<int4> work2;      // two-step list (ELASTIC2, EM2)
  std::vector<int> rows_cols;   // receivers by row (ELASTIC2, EM2)
  std::vector<long> rec_off;    // per shot offset into rows_cols, -1 = none
  long long rows_cols_size = 0; // known before the receivers are read
  std::vector<long> nel, npsi;
};

// Host half of plan creation: validate the inputs and fill the plan's
// settings and everything of `S` except the receivers-by-row index.  Reads
// no device memory.
int stage_host(Plan* p, int kind, int prec, const sdt_model* models, int nmodels,
               const sdt_shot* shots, int nshots, int nt, Stage& S) {
  const bool split = (kind & SDT_SPLIT_CPML) != 0;
  kind &= ~SDT_SPLIT_CPML;
  if (split && kind != SDT_ELASTIC2) return fail(SDT_EINVAL, "SDT_SPLIT_CPML applies to ELASTIC2 only");
  int rc = validate(kind, split, prec, models, nmodels, shots, nshots, nt);
  if (rc) return rc;
  p->kind = kind;
  p->split = split;
  p->prec = prec;
  p->nmodels = nmodels;
  p->nshots = nshots;
  p->nt = nt;
  p->ncomp = kind_3d(kind) ? 3 : 2;
  p->hmodels.assign(models, models + nmodels);
  p->hshots.assign(shots, shots + nshots);
  S.rec_off.assign(nshots, -1);
  for (int s = 0; s < nshots; ++s) {
    p->maxrec = std::max(p->maxrec, shots[s].nrec);
    if (pingpong(kind) && shots[s].nrec > 0) {
      S.rec_off[s] = (long)S.rows_cols_size;
      S.rows_cols_size += models[shots[s].model].nz + 1 + 2LL * shots[s].nrec;
    }
  }
  S.dm.assign(nmodels, DModel());
  for (int m = 0; m < nmodels; ++m) {
    const sdt_model& M = models[m];
    DModel& D = S.dm[m];
    D.nx = M.nx; D.ny = M.ny; D.nz = M.nz;
    D.lx = M.lx; D.ly = kind_3d(kind) ? M.ly : 0; D.lz = M.lz;
    D.dx = M.dx; D.dy = M.dy; D.dz = M.dz; D.dt = M.dt;
    for (int q = 0; q < 4; ++q) D.scal[q] = M.scal[q];
    for (int q = 0; q < 40; ++q) D.coef[q] = M.coef[q];
    for (int a = 0; a < 3; ++a)
      for (int q = 0; q < 6; ++q) D.prof[a][q] = M.prof[a][q];
    D.aux = M.aux;
    p->maxnx = std::max(p->maxnx, M.nx);
    p->maxnz = std::max(p->maxnz, M.nz);
  }
  S.ds.assign(nshots, DShot());
  for (int s = 0; s < nshots; ++s) {
    const sdt_shot& H = shots[s];
    DShot& D = S.ds[s];
    D.model = H.model; D.isrc = H.isrc; D.jsrc = kind_3d(kind) ? H.jsrc : 0; D.ksrc = H.ksrc;
    for (int f = 0; f < 9; ++f) D.f[f] = H.field[f];
    D.psi = H.psi;
    for (int f = 0; f < 9; ++f) D.g[f] = H.field_b[f];
    D.psi2 = H.psi_b;
    for (int c = 0; c < 3; ++c) D.src[c] = H.src[c];
    for (int c = 0; c < 6; ++c) D.msrc[c] = (kind == SDT_ELASTIC2 || kind == SDT_ELASTIC25) ? H.msrc[c] : nullptr;
    D.nrec = H.nrec; D.rec = H.rec; D.trace = H.trace;
    D.blowup = nullptr;
    D.rrow = nullptr;
    D.rxid = nullptr;
  }
  // work lists: (shot, x0, z0); all shots of one tile share blockIdx % 8 (XCD)
  if (kind == SDT_ELASTIC25) S.work = el25_work(p, KB25, EW25 * XW25);
  if (kind == SDT_EM25) S.work = el25_work(p, 1, EW25);  // one plane, one wavefront per item
  if (pingpong(kind)) {
    p->el2_eh = default_rows(p);
    S.work = el2_work(p);
    const int nblk = (int)S.work.size();
    p->el2_grid = std::max(8, std::min(nblk, SDT_EL2_GRID) / 8 * 8);
    int ni = 0;
    S.work2 = build_work2(p, &ni);
    set_work2_grids(p, (int)S.work2.size(), ni);
    // Two steps per launch pay off once the plan has enough tiles to fill the
    // device; a small plan is bound by the latency of one tile's walk, which
    // the two-step kernel lengthens (measured: 4M cells per step and above).
    long long cells = 0;
    for (int s = 0; s < nshots; ++s) cells += (long long)models[shots[s].model].nx * models[shots[s].model].nz;
    p->spl = cells >= SDT_EL2X2_MIN_CELLS ? 2 : 1;
  }
  if (S.work.empty()) S.work.push_back(make_int4(-1, 0, 0, 0));
  p->nwork = (int)S.work.size();
  S.nel.resize(nshots);
  S.npsi.resize(nshots);
  for (int s = 0; s < nshots; ++s) {
    const sdt_model& M = models[shots[s].model];
    S.nel[s] = (long)M.nx * M.ny * M.nz;
    S.npsi[s] = (long)psi_elems(kind | (split ? SDT_SPLIT_CPML : 0), M.nx, M.ny, M.nz, M.lx,
                                kind_3d(kind) ? M.ly : 0, M.lz);
  }
  return SDT_OK;
}

// Read every shot's receivers back from the device (synchronous copies),
// check them against the grid, and build the receivers-by-row index of the
// 2-D kinds: per shot rrow[nz + 1] then (column, index) pairs sorted by
// (row, column).
int stage_receivers(const Plan* p, Stage& S) {
  S.rows_cols.clear();
  S.rows_cols.reserve((size_t)S.rows_cols_size);
  for (int s = 0; s < p->nshots; ++s) {
    const sdt_shot& H = p->hshots[s];
    if (H.nrec <= 0) continue;
    std::vector<int> rec(3 * (size_t)H.nrec);
    if (hipMemcpy(rec.data(), H.rec, rec.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(SDT_EDEVICE, "cannot read receiver indices (not device memory?)");
    const sdt_model& M = p->hmodels[H.model];
    for (int r = 0; r < H.nrec; ++r) {
      const int i = rec[3 * r], j = rec[3 * r + 1], k = rec[3 * r + 2];
      if (i < 0 || i >= M.nx || k < 0 || k >= M.nz || j < 0 || j >= M.ny)
        return fail(SDT_EINVAL, "shot " + std::to_string(s) + ": receiver " + std::to_string(r) +
                                    " outside the padded grid");
    }
    if (S.rec_off[s] < 0) continue;
    std::vector<int> order(H.nrec);
    for (int r = 0; r < H.nrec; ++r) order[r] = r;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      return rec[3 * a + 2] != rec[3 * b + 2] ? rec[3 * a + 2] < rec[3 * b + 2] : rec[3 * a] < rec[3 * b];
    });
    std::vector<int> rrow(M.nz + 1, 0);
    for (int r = 0; r < H.nrec; ++r) rrow[rec[3 * r + 2] + 1]++;
    for (int k = 0; k < M.nz; ++k) rrow[k + 1] += rrow[k];
    S.rows_cols.insert(S.rows_cols.end(), rrow.begin(), rrow.end());
    for (int r : order) {
      S.rows_cols.push_back(rec[3 * r]);
      S.rows_cols.push_back(r);
    }
  }
  return SDT_OK;
}

// Point each shot descriptor at its flag and its part of the receiver index.
void stage_bind(const Plan* p, Stage& S) {
  for (int s = 0; s < p->nshots; ++s) {
    DShot& D = S.ds[s];
    D.blowup = p->dflags + s;
    D.rrow = S.rec_off[s] < 0 ? nullptr : p->drec + S.rec_off[s];
    D.rxid = S.rec_off[s] < 0 ? nullptr : p->drec + S.rec_off[s] + p->hmodels[p->hshots[s].model].nz + 1;
  }
}

// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free run (sdt_run_noalloc): descriptors,
// work lists, receiver index, flags and step counters of one call, carved
// from one static device array per GPU.  Host code never passes `scratch`
// itself to the device: it uses the address hipGetSymbolAddress returns for
// the current device (scratch_ptr).
// ---------------------------------------------------------------------------
#ifndef SDT_NOALLOC_SCRATCH_BYTES
#define SDT_NOALLOC_SCRATCH_BYTES (8u << 20)
#endif
__device__ __align__(256) unsigned char scratch[SDT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (looked up once per device).
int scratch_ptr(unsigned char** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SDT_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(SDT_EDEVICE, "hipGetSymbolAddress returned null for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return SDT_OK;
}

// Offsets of one call's arrays inside the scratch (256-byte aligned).
struct ScratchLayout {
  size_t models, shots, work, work2, rec, flags, ints, nel, npsi, total;
};
ScratchLayout scratch_layout(const Plan* p, const Stage& S) {
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  ScratchLayout L;
  size_t o = 0;
  L.models = o; o = up(o + sizeof(DModel) * p->nmodels);
  L.shots = o;  o = up(o + sizeof(DShot) * p->nshots);
  L.work = o;   o = up(o + sizeof(int4) * S.work.size());
  L.work2 = o;  o = up(o + sizeof(int4) * S.work2.size());
  L.rec = o;    o = up(o + sizeof(int) * (size_t)S.rows_cols_size);
  L.flags = o;  o = up(o + sizeof(int) * p->nshots);
  L.ints = o;   o = up(o + 2 * sizeof(int));
  L.nel = o;    o = up(o + sizeof(long) * p->nshots);
  L.npsi = o;   o = up(o + sizeof(long) * p->nshots);
  L.total = o;
  return L;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* sdt_last_error(void) { return g_err.c_str(); }

const char* sdt_version(void) { return "seidart_hip 0.1 gfx950"; }

int sdt_psi_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz, long long* out) {
  if (!out) return fail(SDT_EINVAL, "out is null");
  const long long n = psi_elems(kind, nx, ny, nz, lx, ly, lz);
  if (n < 0) return fail(SDT_EINVAL, "unknown kind");
  *out = n;
  return SDT_OK;
}

int sdt_plan_create(int kind, int prec, const sdt_model* models, int nmodels,
                    const sdt_shot* shots, int nshots, int nt, void** out_plan) {
  if (!out_plan) return fail(SDT_EINVAL, "out_plan is null");
  *out_plan = nullptr;
  Plan* p = new Plan();
  auto bail = [&](int code) {
    sdt_plan_destroy(p);
    return code;
  };
  Stage S;
  int rc = stage_host(p, kind, prec, models, nmodels, shots, nshots, nt, S);
  if (rc) return bail(rc);
  rc = stage_receivers(p, S);
  if (rc) return bail(rc);
  if (!S.rows_cols.empty()) {
    if (hipMalloc(&p->drec, sizeof(int) * S.rows_cols.size()) != hipSuccess ||
        hipMemcpy(p->drec, S.rows_cols.data(), sizeof(int) * S.rows_cols.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      return bail(fail(SDT_EDEVICE, "hipMalloc/hipMemcpy failed (receivers by row)"));
  }
  if (hipMalloc(&p->dflags, sizeof(int) * nshots) != hipSuccess ||
      hipMalloc(&p->dbase, sizeof(int)) != hipSuccess || hipMalloc(&p->dit0, sizeof(int)) != hipSuccess)
    return bail(fail(SDT_EDEVICE, "hipMalloc failed (flags)"));
  stage_bind(p, S);
  if (hipMalloc(&p->dnel, sizeof(long) * nshots) != hipSuccess ||
      hipMalloc(&p->dnpsi, sizeof(long) * nshots) != hipSuccess ||
      hipMemcpy(p->dnel, S.nel.data(), sizeof(long) * nshots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dnpsi, S.npsi.data(), sizeof(long) * nshots, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SDT_EDEVICE, "hipMalloc/hipMemcpy failed (sizes)"));
  if (hipMalloc(&p->dmodels, sizeof(DModel) * nmodels) != hipSuccess ||
      hipMalloc(&p->dshots, sizeof(DShot) * nshots) != hipSuccess ||
      hipMalloc(&p->dwork, sizeof(int4) * S.work.size()) != hipSuccess)
    return bail(fail(SDT_EDEVICE, "hipMalloc failed (descriptors)"));
  if (hipMemcpy(p->dmodels, S.dm.data(), sizeof(DModel) * nmodels, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dshots, S.ds.data(), sizeof(DShot) * nshots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dwork, S.work.data(), sizeof(int4) * S.work.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(SDT_EDEVICE, "hipMemcpy failed (descriptors)"));
  if (!S.work2.empty()) {
    if (hipMalloc(&p->dwork2, sizeof(int4) * S.work2.size()) != hipSuccess ||
        hipMemcpy(p->dwork2, S.work2.data(), sizeof(int4) * S.work2.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      return bail(fail(SDT_EDEVICE, "hipMalloc/hipMemcpy failed (work list)"));
  }
  hipLaunchKernelGGL(k_fill_int, dim3((nshots + 255) / 256), dim3(256), 0, nullptr, p->dflags,
                     nshots, INT_MAX);
  hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, nullptr, p->dbase, 0, p->dit0, 0);
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(SDT_EDEVICE, "init kernels failed"));
  if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SDT_EDEVICE, "hipStreamCreate failed"));
  *out_plan = p;
  return SDT_OK;
}

int sdt_plan_run(void* plan, int it0, int nsteps, int graph_block, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (it0 < 0 || nsteps < 0 || (long long)it0 + nsteps > p->nt)
    return fail(SDT_EINVAL, "step range outside [0, nt)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  p->last = st;
  if (nsteps == 0) return SDT_OK;
  const bool two = use_two(p, nsteps, graph_block);
  start_of_run(p, st, it0, two);
  hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, st, p->dbase, it0, p->dit0, it0);
  int done = 0;
  if (graph_block > 1 && nsteps >= graph_block) {
    if (p->graph_block != graph_block || !p->gexec || p->graph_two != two) {
      if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
      if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
      HIP_TRY(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
      if (two)
        launch_pairs(p, p->cap, graph_block);
      else
        launch_steps(p, p->cap, graph_block);
      hipLaunchKernelGGL(k_add_int, dim3(1), dim3(1), 0, p->cap, p->dbase, graph_block);
      HIP_TRY(hipStreamEndCapture(p->cap, &p->graph));
      HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
      p->graph_block = graph_block;
      p->graph_two = two;
    }
    for (; done + graph_block <= nsteps; done += graph_block) HIP_TRY(hipGraphLaunch(p->gexec, st));
  }
  if (done < nsteps) {
    if (two)
      launch_pairs(p, st, nsteps - done);
    else
      launch_steps(p, st, nsteps - done);
  }
  end_of_run(p, st, it0 + nsteps, two ? ((nsteps / 2 + (nsteps & 1)) & 1) : -1);
  HIP_TRY(hipGetLastError());
  return SDT_OK;
}

int sdt_plan_profile(void* plan, int it0, int nsteps, void* stream, double* out_ms) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !out_ms) return fail(SDT_EINVAL, "plan or out_ms is null");
  if (it0 < 0 || nsteps <= 0 || (long long)it0 + nsteps > p->nt)
    return fail(SDT_EINVAL, "step range outside [0, nt)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  p->last = st;
  hipEvent_t ev[4];
  for (int q = 0; q < 4; ++q) HIP_TRY(hipEventCreate(&ev[q]));
  double acc[3] = {0, 0, 0};
  const bool two = use_two(p, nsteps, 0);
  start_of_run(p, st, it0, two);
  hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, st, p->dbase, it0, p->dit0, it0);
  int rc = SDT_OK;
  if (two) {
    // two-step launches: slot 0 is the mean per launch (two steps); the
    // single step that ends an odd run is not timed
    int npair = 0;
    for (int s = 0; s < nsteps && rc == SDT_OK; s += 2) {
      const bool single = s + 1 == nsteps;
      if (hipEventRecord(ev[0], st) != hipSuccess) { rc = fail(SDT_EDEVICE, "hipEventRecord"); break; }
      if (p->prec == SDT_FP64)
        launch_two<double>(p, st, s, single);
      else
        launch_two<float>(p, st, s, single);
      if (hipEventRecord(ev[1], st) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess) {
        rc = fail(SDT_EDEVICE, "event sync failed");
        break;
      }
      if (single) break;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) { rc = fail(SDT_EDEVICE, "elapsed"); break; }
      acc[0] += ms;
      ++npair;
    }
    if (rc == SDT_OK) end_of_run(p, st, it0 + nsteps, (nsteps / 2 + (nsteps & 1)) & 1);
    for (int q = 0; q < 4; ++q) (void)hipEventDestroy(ev[q]);
    if (rc) return rc;
    out_ms[0] = npair ? acc[0] / npair : 0.0;
    out_ms[1] = 0.0;
    out_ms[2] = 0.0;
    HIP_TRY(hipGetLastError());
    return SDT_OK;
  }
  for (int s = 0; s < nsteps && rc == SDT_OK; ++s) {
    for (int w = 0; w < 3; ++w) {
      if (hipEventRecord(ev[w], st) != hipSuccess) { rc = fail(SDT_EDEVICE, "hipEventRecord"); break; }
      if (p->prec == SDT_FP64)
        launch_one<double>(p, st, s, w);
      else
        launch_one<float>(p, st, s, w);
    }
    if (rc) break;
    if (hipEventRecord(ev[3], st) != hipSuccess || hipEventSynchronize(ev[3]) != hipSuccess) {
      rc = fail(SDT_EDEVICE, "event sync failed");
      break;
    }
    for (int w = 0; w < 3; ++w) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[w], ev[w + 1]) != hipSuccess) { rc = fail(SDT_EDEVICE, "elapsed"); break; }
      acc[w] += ms;
    }
  }
  if (rc == SDT_OK) end_of_run(p, st, it0 + nsteps);
  for (int q = 0; q < 4; ++q) (void)hipEventDestroy(ev[q]);
  if (rc) return rc;
  for (int w = 0; w < 3; ++w) out_ms[w] = acc[w] / nsteps;
  HIP_TRY(hipGetLastError());
  return SDT_OK;
}

int sdt_plan_check(void* plan, int* first_step, int* per_shot) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  HIP_TRY(hipStreamSynchronize(p->last));
  std::vector<int> f(p->nshots);
  HIP_TRY(hipMemcpy(f.data(), p->dflags, sizeof(int) * p->nshots, hipMemcpyDeviceToHost));
  int first = INT_MAX;
  for (int s = 0; s < p->nshots; ++s) {
    first = std::min(first, f[s]);
    if (per_shot) per_shot[s] = f[s] == INT_MAX ? -1 : f[s];
  }
  if (first_step) *first_step = first == INT_MAX ? -1 : first;
  if (first != INT_MAX)
    return fail(SDT_EUNSTABLE, "code became unstable and blew up at step " + std::to_string(first));
  return SDT_OK;
}

int sdt_run_noalloc(int kind, int prec, const sdt_model* models, int nmodels,
                    const sdt_shot* shots, int nshots, int nt, int it0, int nsteps, void* stream,
                    int* first_step, int* per_shot) {
  if (first_step) *first_step = -1;
  if (it0 < 0 || nsteps < 0 || (long long)it0 + nsteps > nt)
    return fail(SDT_EINVAL, "step range outside [0, nt)");
  // The plan lives on this stack frame; its device arrays point into the
  // scratch, so nothing of it is freed (sdt_plan_destroy is never called).
  Plan P;
  Plan* p = &P;
  Stage S;
  int rc = stage_host(p, kind, prec, models, nmodels, shots, nshots, nt, S);
  if (rc) return rc;
  const ScratchLayout L = scratch_layout(p, S);
  if (L.total > (size_t)SDT_NOALLOC_SCRATCH_BYTES)
    return fail(SDT_EINVAL, "sdt_run_noalloc: this run needs " + std::to_string(L.total) +
                                " bytes of scratch (descriptors, work lists, receiver index) but the "
                                "library holds " + std::to_string((size_t)SDT_NOALLOC_SCRATCH_BYTES) +
                                "; use sdt_plan_create / sdt_plan_run");
  unsigned char* base = nullptr;
  rc = scratch_ptr(&base);
  if (rc) return rc;
  rc = stage_receivers(p, S);
  if (rc) return rc;
  p->dmodels = reinterpret_cast<DModel*>(base + L.models);
  p->dshots = reinterpret_cast<DShot*>(base + L.shots);
  p->dwork = reinterpret_cast<int4*>(base + L.work);
  p->dwork2 = S.work2.empty() ? nullptr : reinterpret_cast<int4*>(base + L.work2);
  p->drec = S.rows_cols.empty() ? nullptr : reinterpret_cast<int*>(base + L.rec);
  p->dflags = reinterpret_cast<int*>(base + L.flags);
  p->dbase = reinterpret_cast<int*>(base + L.ints);
  p->dit0 = p->dbase + 1;
  p->dnel = reinterpret_cast<long*>(base + L.nel);
  p->dnpsi = reinterpret_cast<long*>(base + L.npsi);
  stage_bind(p, S);
  hipStream_t st = static_cast<hipStream_t>(stream);
  p->last = st;
  // the host vectors outlive the copies: the call ends with a stream sync
  HIP_TRY(hipMemcpyAsync(p->dmodels, S.dm.data(), sizeof(DModel) * nmodels, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(p->dshots, S.ds.data(), sizeof(DShot) * nshots, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(p->dwork, S.work.data(), sizeof(int4) * S.work.size(), hipMemcpyHostToDevice, st));
  if (p->dwork2)
    HIP_TRY(hipMemcpyAsync(p->dwork2, S.work2.data(), sizeof(int4) * S.work2.size(),
                           hipMemcpyHostToDevice, st));
  if (p->drec)
    HIP_TRY(hipMemcpyAsync(p->drec, S.rows_cols.data(), sizeof(int) * S.rows_cols.size(),
                           hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(p->dnel, S.nel.data(), sizeof(long) * nshots, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(p->dnpsi, S.npsi.data(), sizeof(long) * nshots, hipMemcpyHostToDevice, st));
  // flags and step counters start fresh every call (the scratch keeps the
  // previous call's contents); cleared by kernels, not hipMemsetAsync
  hipLaunchKernelGGL(k_fill_int, dim3((nshots + 255) / 256), dim3(256), 0, st, p->dflags, nshots, INT_MAX);
  hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, st, p->dbase, it0, p->dit0, it0);
  if (nsteps > 0) {
    const bool two = use_two(p, nsteps, 0);
    start_of_run(p, st, it0, two);
    if (two)
      launch_pairs(p, st, nsteps);
    else
      launch_steps(p, st, nsteps);
    end_of_run(p, st, it0 + nsteps, two ? ((nsteps / 2 + (nsteps & 1)) & 1) : -1);
  }
  HIP_TRY(hipGetLastError());
  std::vector<int> f(nshots);
  HIP_TRY(hipMemcpyAsync(f.data(), p->dflags, sizeof(int) * nshots, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int first = INT_MAX;
  for (int s = 0; s < nshots; ++s) {
    first = std::min(first, f[s]);
    if (per_shot) per_shot[s] = f[s] == INT_MAX ? -1 : f[s];
  }
  if (first_step) *first_step = first == INT_MAX ? -1 : first;
  if (first != INT_MAX)
    return fail(SDT_EUNSTABLE, "code became unstable and blew up at step " + std::to_string(first));
  return SDT_OK;
}

long long sdt_noalloc_scratch_bytes(void) { return (long long)SDT_NOALLOC_SCRATCH_BYTES; }

int sdt_hbm_probe(const void* src, void* dst, long long nfloat, void* stream) {
  if (!src || !dst || nfloat <= 0) return fail(SDT_EINVAL, "sdt_hbm_probe: null buffer or empty size");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_hbm_probe, dim3(256 * 8 * 4), dim3(256), 0, st, static_cast<const float*>(src),
                     static_cast<float*>(dst), nfloat);
  HIP_TRY(hipGetLastError());
  return SDT_OK;
}

// Rebuild the 2-D work lists for a tile shape and order (all or nothing: on
// a failure the plan keeps its tiles, order and work lists), dropping the
// graphs that bake in the old launch shape.
static int retile(Plan* p, int waves_per_block, int rows, int order) {
  // the old work list and the graphs that bake in the launch shape may be in flight
  HIP_TRY(hipDeviceSynchronize());
  const int owb = p->el2_wb, oeh = p->el2_eh, oord = p->el2_order;
  p->el2_wb = waves_per_block;
  p->el2_eh = rows;
  p->el2_order = order;
  const std::vector<int4> work = el2_work(p);
  int4* dw = nullptr;
  int rc = SDT_OK;
  if (hipMalloc(&dw, sizeof(int4) * work.size()) != hipSuccess) {
    dw = nullptr;
    rc = fail(SDT_EDEVICE, "hipMalloc failed (work list)");
  } else if (hipMemcpy(dw, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice) !=
             hipSuccess) {
    rc = fail(SDT_EDEVICE, "hipMemcpy failed (work list)");
  } else if (p->dwork2) {
    rc = upload_work2(p);  // replaces the two-step list only when it succeeds
  }
  if (rc != SDT_OK) {
    if (dw) (void)hipFree(dw);
    p->el2_wb = owb;
    p->el2_eh = oeh;
    p->el2_order = oord;
    return rc;
  }
  (void)hipFree(p->dwork);
  p->dwork = dw;
  p->nwork = (int)work.size();
  p->el2_grid = std::max(8, std::min(p->nwork, SDT_EL2_GRID) / 8 * 8);
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  drop_adjoint_graphs(p);
  return SDT_OK;
}

int sdt_plan_set_tiles(void* plan, int waves_per_block, int rows) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (!pingpong(p->kind)) return fail(SDT_EINVAL, "tile shapes apply to ELASTIC2 and EM2 plans");
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
    return fail(SDT_EINVAL, "waves_per_block must be 1, 2 or 4");
  if (rows < 1 || rows > 256) return fail(SDT_EINVAL, "rows must be in [1, 256]");
  if (waves_per_block == p->el2_wb && rows == p->el2_eh) return SDT_OK;
  return retile(p, waves_per_block, rows, p->el2_order);
}

int sdt_plan_set_tile_order(void* plan, int order) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (!pingpong(p->kind)) return fail(SDT_EINVAL, "tile orders apply to ELASTIC2 and EM2 plans");
  if (order != 0 && order != 1) return fail(SDT_EINVAL, "tile order must be 0 or 1");
  if (order == p->el2_order) return SDT_OK;
  return retile(p, p->el2_wb, p->el2_eh, order);
}

int sdt_plan_get_tile_order(void* plan, int* order) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !order) return fail(SDT_EINVAL, "null argument");
  *order = p->el2_order;
  return SDT_OK;
}

int sdt_plan_set_steps_per_launch(void* plan, int steps) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (steps != 1 && steps != 2) return fail(SDT_EINVAL, "steps per launch must be 1 or 2");
  if (steps == 2 && p->kind != SDT_ELASTIC2 && p->kind != SDT_EM2)
    return fail(SDT_EINVAL, "two steps per launch apply to ELASTIC2 and EM2 plans");
  p->spl = steps;
  return SDT_OK;
}

int sdt_plan_get_steps_per_launch(void* plan, int* steps) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !steps) return fail(SDT_EINVAL, "null argument");
  *steps = p->spl;
  return SDT_OK;
}

int sdt_plan_get_tiles(void* plan, int* waves_per_block, int* rows) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !waves_per_block || !rows) return fail(SDT_EINVAL, "null argument");
  *waves_per_block = p->el2_wb;
  *rows = p->el2_eh;
  return SDT_OK;
}

int sdt_state_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz, long long* out) {
  if (!out) return fail(SDT_EINVAL, "out is null");
  const long long n = state_elems(kind, nx, ny, nz, lx, ly, lz);
  if (n < 0) return fail(SDT_EINVAL, "unknown kind");
  *out = n;
  return SDT_OK;
}

int sdt_adjoint_work_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz,
                           long long* out) {
  (void)ny; (void)lx; (void)ly; (void)lz;
  if (!out) return fail(SDT_EINVAL, "out is null");
  const int base = kind & ~SDT_SPLIT_CPML;
  if (base != SDT_ELASTIC2 && base != SDT_EM2)
    return fail(SDT_EINVAL, "the adjoint is available for ELASTIC2 and EM2 only");
  *out = (base == SDT_ELASTIC2 ? 12LL : 4LL) * nx * nz;
  return SDT_OK;
}

int sdt_adjoint_setup(void* plan, const sdt_adj_shot* adj, int nshots, unsigned grad_mask,
                      int flags) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !adj) return fail(SDT_EINVAL, "plan or adj is null");
  if (grad_mask & ~0xfffu) return fail(SDT_EINVAL, "grad_mask has bits beyond the 12 slots");
  if (flags & ~SDT_ADJ_REFERENCE) return fail(SDT_EINVAL, "unknown adjoint flags");
  if (p->kind != SDT_ELASTIC2 && p->kind != SDT_EM2)
    return fail(SDT_EINVAL, "the adjoint is available for ELASTIC2 and EM2 only");
  if (nshots != p->nshots) return fail(SDT_EINVAL, "adj must hold one entry per shot of the plan");
  if (p->kind == SDT_EM2 && (grad_mask & ~0xfu))
    return fail(SDT_EINVAL, "EM2 has 4 coefficient slots (grad_mask <= 0xf)");
  std::vector<DAdj> h(nshots);
  for (int s = 0; s < nshots; ++s) {
    const sdt_adj_shot& a = adj[s];
    const std::string tag = "shot " + std::to_string(s) + ": ";
    if (!a.trail[0] || !a.lam[0] || !a.lam[1] || !a.work || !a.grad)
      return fail(SDT_EINVAL, tag + "trail[0], lam, work and grad must be set");
    if ((a.trail[1] != nullptr) != (adj[0].trail[1] != nullptr))
      return fail(SDT_EINVAL, tag + "trail[1] must be set for every shot or for none");
    if (a.trail_len < 2) return fail(SDT_EINVAL, tag + "trail_len must be at least 2");
    const sdt_model& M = p->hmodels[p->hshots[s].model];
    DAdj& D = h[s];
    if (p->kind == SDT_ELASTIC2 || p->kind == SDT_EM2) {
      // the fused kernels address a state and a gradient with 32-bit byte offsets
      const long long es = p->prec == SDT_FP64 ? 8 : 4;
      const long long st = state_elems(p->kind | (p->split ? SDT_SPLIT_CPML : 0), M.nx, M.ny, M.nz,
                                       M.lx, M.ly, M.lz);
      if (st * es >= (1LL << 32) || 12LL * M.nx * M.nz * es >= (1LL << 32))
        return fail(SDT_EINVAL, tag + "grid too large for the adjoint (state or gradient over 4 GiB)");
    }
    D.trail[0] = a.trail[0];
    D.trail[1] = a.trail[1];
    D.slot = (long)state_elems(p->kind | (p->split ? SDT_SPLIT_CPML : 0), M.nx, M.ny, M.nz, M.lx,
                               kind_3d(p->kind) ? M.ly : 0, M.lz);
    D.trail_len = a.trail_len;
    D.lam[0] = a.lam[0];
    D.lam[1] = a.lam[1];
    D.work = a.work;
    D.resid = a.resid;
    D.grad = a.grad;
    D.gsrc = a.gsrc;
  }
  // the previous descriptors may still be read by queued kernels
  HIP_TRY(hipDeviceSynchronize());
  if (!p->dadj) HIP_TRY(hipMalloc(&p->dadj, sizeof(DAdj) * nshots));
  p->gmask = grad_mask;
  p->adj_flags = flags;
  drop_adjoint_graphs(p);  // they bake in the options
  // FP32 staging of the gradient, one per distinct grad array (fused kernel)
  for (float* g : p->g32) (void)hipFree(g);
  p->g32.clear();
  if (p->dflush) { (void)hipFree(p->dflush); p->dflush = nullptr; }
  p->nflush = 0;
  // (the EM2 reference kernels stage in FP32 too; the elastic ones add in FP64)
  if (p->prec == SDT_FP32 && (p->kind == SDT_EM2 || !(flags & SDT_ADJ_REFERENCE))) {
    std::vector<DFlush> fl;
    for (int s = 0; s < nshots; ++s) {
      const sdt_model& M = p->hmodels[p->hshots[s].model];
      const long nel = (long)M.nx * M.nz;
      int f = 0;
      while (f < (int)fl.size() && fl[f].grad != h[s].grad) ++f;
      if (f == (int)fl.size()) {
        float* g = nullptr;
        HIP_TRY(hipMalloc(&g, sizeof(float) * 12 * nel));
        HIP_TRY(hipMemset(g, 0, sizeof(float) * 12 * nel));
        p->g32.push_back(g);
        fl.push_back(DFlush{h[s].grad, g, nel});
      } else if (fl[f].n != nel) {
        return fail(SDT_EINVAL, "shots sharing a grad array must have grids of one size");
      }
      h[s].g32 = fl[f].g32;
    }
    HIP_TRY(hipMalloc(&p->dflush, sizeof(DFlush) * fl.size()));
    HIP_TRY(hipMemcpy(p->dflush, fl.data(), sizeof(DFlush) * fl.size(), hipMemcpyHostToDevice));
    p->nflush = (int)fl.size();
  }
  HIP_TRY(hipMemcpy(p->dadj, h.data(), sizeof(DAdj) * nshots, hipMemcpyHostToDevice));
  p->hadj = h;
  if (!p->dtbase) HIP_TRY(hipMalloc(&p->dtbase, sizeof(int)));
  if (!p->dtsel_trail) HIP_TRY(hipMalloc(&p->dtsel_trail, sizeof(int)));
  if (!p->dtsel_adj) HIP_TRY(hipMalloc(&p->dtsel_adj, sizeof(int)));
  if (!p->dadjwork) {
    // (shot, x0, z0, rows) items of AWB * AW columns x rows; the items of one
    // tile over all shots of its model share blockIdx % 8 (one XCD's L2).
    // The elastic kernel walks AH rows; the EM kernel takes the tallest of
    // AH, 16, 8, 4, 2 rows that gives the plan 1000 tiles or more (its
    // single-shot grids are small and stay in the cache, so the walk's
    // latency is what a tile costs).
    const int bw = AW * AWB;
    int eh = AH;
    if (p->kind == SDT_EM2) {
      for (int rows : {AH, 16, 8, 4, 2}) {
        eh = rows;
        long long tiles = 0;
        for (const sdt_shot& S : p->hshots) {
          const sdt_model& M = p->hmodels[S.model];
          tiles += (long long)((M.nx + bw - 1) / bw) * ((M.nz + rows - 1) / rows);
        }
        if (tiles >= 1000) break;
      }
    }
    std::vector<int4> work;
    for (int m = 0; m < p->nmodels; ++m) {
      const sdt_model& M = p->hmodels[m];
      const int tx = (M.nx + bw - 1) / bw, tz = (M.nz + eh - 1) / eh, ntile = tx * tz;
      std::vector<int> ms;
      for (int s = 0; s < p->nshots; ++s)
        if (p->hshots[s].model == m) ms.push_back(s);
      if (ms.empty()) continue;
      for (int tg = 0; tg < (ntile + 7) / 8; ++tg)
        for (int s : ms)
          for (int tt = 0; tt < 8; ++tt) {
            const int t = tg * 8 + tt;
            work.push_back(t < ntile ? make_int4(s, (t % tx) * bw, (t / tx) * eh, eh)
                                     : make_int4(-1, 0, 0, 0));
          }
    }
    HIP_TRY(hipMalloc(&p->dadjwork, sizeof(int4) * work.size()));
    HIP_TRY(hipMemcpy(p->dadjwork, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice));
    p->nadjwork = (int)work.size();
  }

  return SDT_OK;
}

int sdt_plan_run_trail(void* plan, int it0, int nsteps, int buf, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (!p->dadj) return fail(SDT_EINVAL, "call sdt_adjoint_setup first");
  if (buf < 0 || buf > 1 || !p->hadj[0].trail[buf]) return fail(SDT_EINVAL, "no such trail buffer");
  if (it0 < 0 || nsteps < 0 || (long long)it0 + nsteps > p->nt)
    return fail(SDT_EINVAL, "step range outside [0, nt)");
  for (const DAdj& a : p->hadj)
    if (nsteps + 1 > a.trail_len) return fail(SDT_EINVAL, "trail too short for nsteps + 1 states");
  hipStream_t st = static_cast<hipStream_t>(stream);
  p->last = st;
  if (nsteps == 0) return SDT_OK;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p->dbase, it0);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p->dtsel_trail, buf);
  if (nsteps < 2) {
    launch_trail(p, st, nsteps);
  } else {
    const int rc = replay_segment(p, 0, nsteps, st, [&](hipStream_t c) { launch_trail(p, c, nsteps); });
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return SDT_OK;
}

int sdt_adjoint_run(void* plan, int t_hi, int t_lo, int trail_t0, int buf, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return fail(SDT_EINVAL, "plan is null");
  if (!p->dadj) return fail(SDT_EINVAL, "call sdt_adjoint_setup first");
  if (buf < 0 || buf > 1 || !p->hadj[0].trail[buf]) return fail(SDT_EINVAL, "no such trail buffer");
  if (t_lo < 0 || t_hi < t_lo || t_hi > p->nt || trail_t0 < 0 || trail_t0 > t_lo)
    return fail(SDT_EINVAL, "need 0 <= trail_t0 <= t_lo <= t_hi <= nt");
  for (const DAdj& a : p->hadj)
    if (t_hi - trail_t0 + 1 > a.trail_len)
      return fail(SDT_EINVAL, "trail does not hold the states up to t_hi");
  hipStream_t st = static_cast<hipStream_t>(stream);
  p->last = st;
  if (t_hi == t_lo) return SDT_OK;
  const int khi = t_hi - trail_t0, klo = t_lo - trail_t0;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p->dtbase, trail_t0);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p->dtsel_adj, buf);
  if (khi - klo < 2) {
    launch_adjoint(p, st, khi, klo);
  } else {
    const long long key = (long long)khi << 32 | (unsigned)klo;
    const int rc = replay_segment(p, 1, key, st, [&](hipStream_t c) { launch_adjoint(p, c, khi, klo); });
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return SDT_OK;
}

void sdt_plan_destroy(void* plan) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return;
  (void)hipDeviceSynchronize();  // nothing of the plan may still be in flight
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->dmodels) (void)hipFree(p->dmodels);
  if (p->dshots) (void)hipFree(p->dshots);
  if (p->dwork) (void)hipFree(p->dwork);
  if (p->dwork2) (void)hipFree(p->dwork2);
  if (p->dit0) (void)hipFree(p->dit0);
  if (p->drec) (void)hipFree(p->drec);
  if (p->dflags) (void)hipFree(p->dflags);
  if (p->dbase) (void)hipFree(p->dbase);
  if (p->dnel) (void)hipFree(p->dnel);
  if (p->dnpsi) (void)hipFree(p->dnpsi);
  if (p->dadj) (void)hipFree(p->dadj);
  if (p->dadjwork) (void)hipFree(p->dadjwork);
  if (p->dtbase) (void)hipFree(p->dtbase);
  if (p->dtsel_trail) (void)hipFree(p->dtsel_trail);
  if (p->dtsel_adj) (void)hipFree(p->dtsel_adj);
  drop_adjoint_graphs(p);
  for (float* g : p->g32) (void)hipFree(g);
  if (p->dflush) (void)hipFree(p->dflush);
  delete p;
}

}  // extern "C"
