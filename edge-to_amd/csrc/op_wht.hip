// Fine-grid H8 operator in the corner Walsh-Hadamard basis (FP32 or FP64
// vector ALU), element-centric, marching in y.
//
// Algebra.  Number the corners of a brick element by their offset bits,
// beta = hx + 2*hy + 4*hz, and let H be the 8-point Walsh-Hadamard matrix,
// H[M][beta] = (-1)^popcount(M & beta).  The reflections of the brick (a
// corner map beta -> beta ^ g together with a sign flip of the reflected
// displacement components) leave K0 invariant, so in the transformed
// coordinates  uh_c = H u_c  (per displacement component c) K0 is block
// diagonal:
//     yh_c[M] = sum_d G[M][c][d] * uh_d[M ^ e_c ^ e_d],   y_c = H yh_c,
// with e_c = 1 << c and G[M][c][d] = (1/64) sum H[M][b] K0[(b,c),(b',d)]
// H[M^e_c^e_d][b'].  G has 45 non-zero entries (block M = 0, the rigid
// translations, vanishes; block 7 is diagonal; every other block misses the
// pair of axes given by its bits or their complement), so an element product
// costs 45 FMAs plus two separable transforms instead of 576 FMAs.
//
// The transforms are split into the x-z face stage (4-point, per node face)
// and the y stage (sum / difference of the bottom and top faces).  Marching
// in y, the top face of layer ey is the bottom face of layer ey+1: its
// forward face transform and its transformed output are carried in
// registers, so each element transforms one new face and finishes one face
// per layer.
//
// Assembly without atomics, in a fixed order.  Thread = element; a
// wavefront holds 64 elements consecutive along z.  The hz = 1 corner
// contributions move one lane up (DPP wave_shr:1; lane 63's value goes to a
// seam slot in LDS) and are added to the hz = 0 ones, giving per element the
// two node contributions (hx = 0, 1) of node k = ez.  Those are written to
// LDS and the node phase adds the contributions of element columns i and
// i-1 (and the seam slot at k = 64, 128).
//
// Coefficients.  G's 45 non-zeros depend on K0 only: the host builds them once
// (wht_table) and every launch takes them as a by-value kernel argument, which
// the contraction reads with scalar loads per output component (held in scalar
// registers for the whole kernel they do not fit next to the other scalars and
// spill to vector lanes).
//
// Tiling.  A workgroup (8 wavefronts) owns TC element columns in x (16 for
// nelz <= 64, 8 for nelz <= 128; all z), completing TC-1 node columns, and
// marches in y; the (x tile, node plane) units are split evenly over a
// persistent grid, each range starting with one extra layer that seeds the
// carries.  LDS holds two staged node planes (SoA, constrained DOFs zeroed),
// one layer of E and the per-column node contributions; the next plane and
// layer are loaded into registers while the current layer is computed.
#include "common.hpp"
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

constexpr int kWG = 63;  // transformed blocks M = 1..7, 3 x 3 each

// Workgroup of NT threads on a tile of NS slices of 64 elements per layer
// (NS = 16: TC = 16 element columns for nelz <= 64, 8 for nelz <= 128; NS = 8:
// half of that).  Per-workgroup capacities: (TC+1)*nz <= 96 NS staged nodes,
// TC*nelz <= 64 NS E values, (TC-1)*nz <= 64 NS output nodes.
template <int NT, int NS>
struct WCfg {
  static constexpr int waves = NT / 64;
  static constexpr int stage_cap = 96 * NS, e_cap = 64 * NS, node_cap = 64 * NS;
  static constexpr int slots = NS / waves;                   // elements per thread per layer
  static constexpr int stage = (stage_cap + NT - 1) / NT;    // staged nodes per thread
  static constexpr int estage = e_cap / NT;                  // staged E per thread
  static constexpr int nodes = node_cap / NT;                // output nodes per thread
  static_assert(slots >= 1 && estage >= 1 && nodes >= 1, "tile smaller than the workgroup");
};

// structural non-zeros of G[M][c][d] (M = 1..7)
__host__ __device__ constexpr bool wht_nz(int M, int c, int d) {
  if (M == 0) return false;
  if (c == d) return true;
  if (M == 7) return false;
  const int pair = (1 << c) | (1 << d);
  const int bits = __builtin_popcount(M);
  // single-bit M: zero on the pair of the other two axes; two-bit M: zero on its own pair
  return bits == 1 ? (pair & M) != 0 : pair != M;
}

__host__ inline int wht_tc(const Grid& g, int ns) { return g.ez <= 64 ? ns : ns / 2; }

__host__ inline size_t wht_lds_values(const Grid& g, int ns) {
  const int tc = wht_tc(g, ns);
  const size_t nz = g.nz(), span = (size_t)(tc + 1) * nz;
  return 6 * span                 // two staged node planes
         + 6 * (size_t)tc * nz    // node contributions [3][2][tc][nz]
         + 6 * (size_t)tc * 2     // seam slots [3][2][tc][2]
         + (size_t)tc * g.ez;     // E of one layer
}

// G's non-zeros in contraction order, 16 per output component (the last one 0);
// the kernel's first argument, so that it sits at offset 0 of the kernel argument
// segment
template <typename T>
struct WhtK {
  T g[48];
};
static_assert(sizeof(WhtK<double>) == 384 && sizeof(WhtK<float>) == 192, "coefficient table layout");

void wht_table(const double* K0, WhtTable* t) {
  double G[kWG];
  for (int q = 0; q < kWG; ++q) {
    const int M = 1 + q / 9, c = (q / 3) % 3, d = q % 3;
    const int Mp = M ^ (1 << c) ^ (1 << d);
    double s = 0.0;
    for (int bb = 0; bb < 8; ++bb) {
      const int a = h8_corner(bb & 1, (bb >> 1) & 1, bb >> 2);
      for (int bp = 0; bp < 8; ++bp) {
        const int ap = h8_corner(bp & 1, (bp >> 1) & 1, bp >> 2);
        const int sg1 = __builtin_popcount(M & bb) & 1, sg2 = __builtin_popcount(Mp & bp) & 1;
        const double k = K0[(3 * a + c) * 24 + 3 * ap + d];
        s += (sg1 ^ sg2) ? -k : k;
      }
    }
    G[q] = s / 64.0;
  }
  // G is symmetric, G[M][c][d] = G[M ^ e_c ^ e_d][d][c]: the entries with c <= d are used
  for (int c = 0; c < 3; ++c) {
    int i = 0;
    for (int M = 1; M < 8; ++M)
      for (int d = 0; d < 3; ++d)
        if (wht_nz(M, c, d))
          t->g[16 * c + i++] = c <= d ? G[(M - 1) * 9 + c * 3 + d] : G[((M ^ (1 << c) ^ (1 << d)) - 1) * 9 + d * 3 + c];
    for (; i < 16; ++i) t->g[16 * c + i] = 0.0;
  }
}

// lane j takes lane j-1 of the wavefront, lane 0 takes 0
__device__ inline unsigned wht_shr32(unsigned v) { return __builtin_amdgcn_update_dpp(0u, v, 0x138, 0xf, 0xf, false); }
__device__ inline float wht_shr(float v) {
  return __builtin_bit_cast(float, wht_shr32(__builtin_bit_cast(unsigned, v)));
}
__device__ inline double wht_shr(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = wht_shr32((unsigned)u), hi = wht_shr32((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 4-point transform over (hx, hz): F[px + 2 pz] = sum (-1)^(px hx + pz hz) u[hx][hz]
template <typename T>
__device__ inline void wht_face(const T u00, const T u10, const T u01, const T u11, T* F) {
  const T a0 = u00 + u10, a1 = u00 - u10, b0 = u01 + u11, b1 = u01 - u11;
  F[0] = a0 + b0;
  F[1] = a1 + b1;
  F[2] = a0 - b0;
  F[3] = a1 - b1;
}

// FP32 on 8-slice tiles of 512 threads runs two workgroups per compute unit:
// at most 128 VGPRs (4 waves per SIMD)
template <typename T, int NT, int NS>
constexpr int wht_min_waves() {
  return (sizeof(T) == 4 && NT == 512 && NS == 8) ? 4 : 1;
}

template <typename T, int MODE, int NT, int NS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(wht_min_waves<T, NT, NS>())))
void op0_wht_kernel(const WhtK<T> Gk, const MxArgs A,
                                                            const T* __restrict__ E,
                                                            const uint8_t* __restrict__ cons,
                                                            const T* __restrict__ x,
                                                            const T* __restrict__ b,
                                                            const T* __restrict__ Dinv,
                                                            const double* __restrict__ omega,
                                                            double wscale, const T* __restrict__ xadd,
                                                            T* __restrict__ y, double* __restrict__ partial,
                                                            const T* __restrict__ dotv) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr bool kAdd = MODE == TM_JAC_ADD || MODE == TM_JAC_ADD_DOT;  // output + xadd
  constexpr bool kJac = MODE == TM_JAC || MODE == TM_JAC0 || kAdd;
  constexpr bool kDot = MODE == TM_YDOT || MODE == TM_JAC_ADD_DOT;  // per-workgroup partial sums
  constexpr bool kZero = MODE == TM_JAC0;  // input x0 = w D^-1 b, built while staging
  constexpr bool kProd = MODE == TM_Y || MODE == TM_YDOT;  // output is A x
  // kEarly: the output plane's node operands are requested before the next
  // layer's staging loads, and the staging loads are unconditional (slots past
  // the span and planes past the range read index 0 and are not stored), so the
  // compiler can count the loads in flight and the node phase waits only for
  // its own operands.  Measured per mode at 256x128x128: faster for the FP64
  // modes and the FP32 residual and zero-guess sweeps, slower for the FP32
  // Jacobi sweeps that carry more node operands.
  constexpr bool kEarly = sizeof(T) == 8 || !(MODE == TM_JAC || kAdd);
  const Grid& g = A.g;
  const int nx = g.nx(), ny = g.ny(), nz = g.nz();
  const int nelx = g.ex, nely = g.ey, nelz = g.ez;
  const int spc = nelz <= 64 ? 1 : 2;  // 64-element slices per element column
  const int tc = NS / spc;             // element columns per x tile
  const int span = (tc + 1) * nz;
  T* su = reinterpret_cast<T*>(lds_raw);  // [2 slots][3 comps][span]
  T* sw = su + 6 * span;                  // [3 comps][2 hx][tc][nz]
  T* ss = sw + 6 * tc * nz;               // seams [3 comps][2 hx][tc][2]
  T* se = ss + 12 * tc;                   // [tc][nelz]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  // 32-bit DOF indices and byte offsets: the host checks 3 * n_nodes < 2^31 and 3 * n_nodes * sizeof(T) < 2^32
  const int plane3 = nx * nz * 3;
  // Without constraints every constraint-byte read goes to byte 0 of E and is
  // masked to zero.  Both candidates are kernel arguments, so the loads stay
  // global loads (a pointer that may also point at a device-side constant
  // becomes a flat load, which also counts against the LDS wait counter).
  const uint8_t* cp = cons != nullptr ? cons : reinterpret_cast<const uint8_t*>(E);
  const int cmask = cons != nullptr ? ~0 : 0;
  const uint8_t cbyte = cons != nullptr ? (uint8_t)0xff : (uint8_t)0;
  constexpr int kWThreads = NT;
  using Cfg = WCfg<NT, NS>;
  constexpr int kWSlots = Cfg::slots;
  constexpr int kWStage = Cfg::stage;
  constexpr int kWEStage = Cfg::estage;
  constexpr int kWNodes = Cfg::nodes;

  const T w = kJac ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  double dot = 0.0;  // this thread's share of x.y (TM_YDOT) or dotv.y (TM_JAC_ADD_DOT)

  // element slots: slice s = wave + 8 m -> column s / spc, z range 64 (s % spc) + [0, 64)
  int scl[kWSlots], sez[kWSlots];
#pragma unroll
  for (int m = 0; m < kWSlots; ++m) {
    const int s = wave + Cfg::waves * m;
    scl[m] = s / spc;
    sez[m] = 64 * (s % spc) + lane;
  }
  // staged-node and E slots of this thread
  int snl[kWStage], sk[kWStage], ecl[kWEStage], eez[kWEStage];
#pragma unroll
  for (int m = 0; m < kWStage; ++m) {
    const int s = tid + kWThreads * m;
    snl[m] = s / nz;
    sk[m] = s - snl[m] * nz;
  }
#pragma unroll
  for (int m = 0; m < kWEStage; ++m) {
    const int t = tid + kWThreads * m;
    ecl[m] = t / nelz;
    eez[m] = t - ecl[m] * nelz;
  }

  // output nodes of this thread: staged slot s = nz + tid + 512 q (output columns start at 1)
  int osl[kWNodes], onl[kWNodes], ok_[kWNodes];
#pragma unroll
  for (int q = 0; q < kWNodes; ++q) {
    osl[q] = nz + tid + kWThreads * q;
    onl[q] = osl[q] / nz;
    ok_[q] = osl[q] - onl[q] * nz;
  }

  const int64_t u0 = (int64_t)blockIdx.x * A.units / gridDim.x;
  const int64_t u1 = (int64_t)(blockIdx.x + 1) * A.units / gridDim.x;
  int64_t un = u0;
  while (un < u1) {
    const int xt = (int)(un / ny);
    const int ja = (int)(un - (int64_t)xt * ny);
    const int jb = (int)min((int64_t)ny, (int64_t)ja + (u1 - un));
    un += jb - ja;
    const int c0 = xt * (tc - 1) - 1;  // element column of local column 0 = node column of staged column 0

    // Staging loads are unconditional (out-of-range slots read index 0) and their
    // masks are applied when the registers are written to LDS, so a layer's loads
    // are all in flight together instead of each waiting on its constraint byte.
    T pu[kWStage][3];
    T pdv[kZero ? kWStage : 1][3];  // TM_JAC0: D^-1 of the staged nodes
    uint8_t pk[kWStage][3];
    bool pv[kWStage];
    T pe[kWEStage];
    bool pev[kWEStage];
    auto fetch_u = [&](int jp) {
      const bool pin = jp >= 0 && jp < ny;
      const int pb = (pin ? jp : 0) * plane3 + c0 * nz * 3;
#pragma unroll
      for (int m = 0; m < kWStage; ++m) {
        const int s = tid + kWThreads * m;
        const int n = c0 + snl[m];
        pv[m] = pin && s < span && n >= 0 && n < nx;
        if (kEarly || m == 0 || kWThreads * m + wave_u * 64 < span) {  // wave-uniform
          const int d = pv[m] ? pb + s * 3 : 0;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            if (kZero) {
              pu[m][r] = ld32(b, d, r);
              pdv[kZero ? m : 0][r] = ld32(Dinv, d, r);
            } else {
              pu[m][r] = ld32(x, d, r);
            }
            pk[m][r] = ld32(cp, d & cmask, r) & cbyte;
          }
        }
      }
    };
    auto store_u = [&](int jp) {
      T* dst = su + (size_t)(jp & 1) * 3 * span;
#pragma unroll
      for (int m = 0; m < kWStage; ++m) {
        const int s = tid + kWThreads * m;
        if (s < span) {
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const T v = kZero ? w * pdv[kZero ? m : 0][r] * pu[m][r] : pu[m][r];
            dst[r * span + s] = (pv[m] && pk[m][r] == 0) ? v : T(0);
          }
        }
      }
    };
    auto fetch_e = [&](int ey) {
      const bool lin = ey >= 0 && ey < nely;
#pragma unroll
      for (int m = 0; m < kWEStage; ++m) {
        const int c = c0 + ecl[m];
        pev[m] = lin && ecl[m] < tc && c >= 0 && c < nelx;
        pe[m] = ld32(E, pev[m] ? c * (int)A.e_sx + ey * (int)A.e_sy + eez[m] : 0);
      }
    };
    auto store_e = [&]() {
#pragma unroll
      for (int m = 0; m < kWEStage; ++m) {
        const int t = tid + kWThreads * m;
        if (t < tc * nelz) se[t] = pev[m] ? pe[m] : T(0);
      }
    };
    // face transform of the face of plane jp (slot jp & 1) for element slot m
    auto face = [&](int jp, int m, T (&F)[3][4]) {
      const int ezc = min(sez[m], nelz - 1);
      const T* ub = su + (size_t)(jp & 1) * 3 * span + scl[m] * nz + ezc;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T* uc = ub + c * span;
        wht_face(uc[0], uc[nz], uc[1], uc[nz + 1], F[c]);
      }
    };

    // prologue: node planes ja-1 and ja, element layer ja-1 (the layer's E is
    // requested with the first plane)
    fetch_u(ja - 1);
    fetch_e(ja - 1);
    store_u(ja - 1);
    store_e();
    fetch_u(ja);
    store_u(ja);
    __syncthreads();

    // FP32 carries the transformed bottom face from the previous layer; FP64
    // (twice the registers) recomputes it from the staged plane.
    constexpr bool kCarryF = sizeof(T) == 4;
    T Fb[kCarryF ? kWSlots : 1][3][4];  // transformed bottom face of the current layer
    T Yc[kWSlots][3][4];                // transformed top-face output of the previous layer
#pragma unroll
    for (int m = 0; m < kWSlots; ++m) {
      if (kCarryF) face(ja - 1, m, Fb[kCarryF ? m : 0]);
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int f = 0; f < 4; ++f) Yc[m][c][f] = T(0);
    }

    for (int ey = ja - 1; ey < jb; ++ey) {
      const bool out = ey >= ja;
      const bool more = ey + 2 <= jb;
      if (!kEarly && more) {
        fetch_u(ey + 2);
        fetch_e(ey + 1);
      }
      // global operands of this layer's output plane, requested before the element phase
      const int jb3 = ey * plane3 + c0 * nz * 3;
      bool act[kWNodes];
      int dd[kWNodes];
      T bq[kWNodes][3], dq[kWNodes][3], aq[kWNodes][3], xq[kWNodes][3];
      T vq[MODE == TM_JAC_ADD_DOT ? kWNodes : 1][3];
      uint8_t cq[kWNodes][3];
#pragma unroll
      for (int q = 0; q < kWNodes; ++q) {
        act[q] = out && onl[q] < tc && c0 + onl[q] < nx;
        dd[q] = act[q] ? jb3 + osl[q] * 3 : 0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          cq[q][r] = ld32(cp, dd[q] & cmask, r) & cbyte;
          bq[q][r] = (!kProd && !kZero) ? ld32(b, dd[q], r) : T(0);
          dq[q][r] = kJac ? ld32(Dinv, dd[q], r) : T(0);
          aq[q][r] = kAdd ? ld32(xadd, dd[q], r) : T(0);
          if (MODE == TM_JAC_ADD_DOT) vq[MODE == TM_JAC_ADD_DOT ? q : 0][r] = ld32(dotv, dd[q], r);
        }
      }
      if (kEarly) {
        fetch_u(ey + 2);
        fetch_e(ey + 1);
      }

      // element phase: layer ey
#pragma unroll
      for (int m = 0; m < kWSlots; ++m) {
        const int cl = scl[m], ez = sez[m];
        T Ft[3][4];
        face(ey + 1, m, Ft);
        T (&Fbm)[3][4] = Fb[kCarryF ? m : 0];
        if (!kCarryF) face(ey, m, Fbm);
        const T e = ez < nelz ? se[cl * nelz + ez] : T(0);
        T uh[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const int M0 = (f & 1) | ((f >> 1) << 2);  // px + 4 pz
            uh[c][M0] = e * (Fbm[c][f] + Ft[c][f]);
            uh[c][M0 | 2] = e * (Fbm[c][f] - Ft[c][f]);
            Fbm[c][f] = Ft[c][f];
          }
        // one output component at a time: its 8 transformed values, the y stage
        // and face transform of the inverse, and its node contributions (so
        // only 8 of the 24 transformed outputs are live at once)
        const int k = ez;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // this component's 15 coefficients: scalar loads from the kernel argument
          // segment (Gk is the first argument, at offset 0; taking Gk's address
          // instead would make a private copy), through a pointer the compiler may
          // not hoist out of the layer loop
          const __attribute__((address_space(4))) T* gq =
              (const __attribute__((address_space(4))) T*)__builtin_amdgcn_kernarg_segment_ptr() + 16 * c;
          asm volatile("" : "+s"(gq));
          T yc[8];
          yc[0] = T(0);
          int gi = 0;
#pragma unroll
          for (int M = 1; M < 8; ++M) {
            T s = T(0);
#pragma unroll
            for (int d = 0; d < 3; ++d)
              if (wht_nz(M, c, d)) s = fma(gq[gi++], uh[d][M ^ (1 << c) ^ (1 << d)], s);
            yc[M] = s;
          }
          // y stage of the inverse; this layer's bottom face completes node plane ey
          T Yb[4];
#pragma unroll
          for (int f = 0; f < 4; ++f) {
            const int M0 = (f & 1) | ((f >> 1) << 2);
            Yb[f] = yc[M0] + yc[M0 | 2] + Yc[m][c][f];
            Yc[m][c][f] = yc[M0] - yc[M0 | 2];
          }
          T o[4];  // [hx + 2 hz]
          wht_face(Yb[0], Yb[1], Yb[2], Yb[3], o);  // symmetric; the normalisation is in G
          // z: the hz = 1 contributions move one lane up; lane 63's go to the seam slot
#pragma unroll
          for (int hx = 0; hx < 2; ++hx) {
            const T up = o[hx + 2];
            const T zv = o[hx] + wht_shr(up);
            if (out && k <= nelz) sw[((c * 2 + hx) * tc + cl) * nz + k] = zv;
            if (out && lane == 63) ss[((c * 2 + hx) * tc + cl) * 2 + (ez >> 6)] = up;
          }
          if (c < 2) __builtin_amdgcn_sched_barrier(0);
        }
      }
      // the node's own (staged, constrained-zeroed) value is read before the
      // barrier: after it, the slot of plane ey is refilled by other threads
      if (MODE != TM_Y) {
        const T* own = su + (size_t)(ey & 1) * 3 * span;
#pragma unroll
        for (int q = 0; q < kWNodes; ++q) {
          const int os = min(osl[q], span - 1);
#pragma unroll
          for (int r = 0; r < 3; ++r) xq[q][r] = own[r * span + os];
        }
      }
      __syncthreads();

      // node phase: node plane ey
      if (out) {
#pragma unroll
        for (int q = 0; q < kWNodes; ++q) {
          if (act[q]) {
            const int nl = onl[q], k = ok_[q];
            const bool wv = k < 64 * spc;  // contribution slot written by a lane
            const bool seam = k > 0 && (k & 63) == 0;
            // both slot kinds are read at clamped (always written) indices and
            // selected afterwards, so the reads do not wait one by one
            const int kw = min(k, 64 * spc - 1);
            const int ks = max((k >> 6) - 1, 0);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const int d = dd[q];
              const T s0 = sw[((r * 2) * tc + nl) * nz + kw] + sw[((r * 2 + 1) * tc + nl - 1) * nz + kw];
              const T s1 = ss[((r * 2) * tc + nl) * 2 + ks] + ss[((r * 2 + 1) * tc + nl - 1) * 2 + ks];
              const T ax0 = (wv ? s0 : T(0)) + (seam ? s1 : T(0));
              const bool cr = cq[q][r] != 0;
              T xv = xq[q][r];  // staged value: zero on constrained DOFs
              T vv;
              // unconstrained row
              if (kZero)
                // x0 = w D^-1 b is the staged value, so x0 + w D^-1 (b - A x0) = 2 x0 - w D^-1 A x0
                vv = xv + xv - w * dq[q][r] * ax0;
              else if (kProd)
                vv = ax0;
              else if (MODE == TM_RES)
                vv = bq[q][r] - ax0;
              else
                vv = xv + w * dq[q][r] * (bq[q][r] - ax0);
              if (cr) {
                // identity row (A x = x): the value is read here, where it is used, so
                // only waves holding a constrained DOF wait for it
                if (kZero) {
                  const T bv = ld32(b, d, r);
                  const T x0 = w * dq[q][r] * bv;
                  vv = x0 + w * dq[q][r] * (bv - x0);
                } else {
                  xv = ld32(x, d, r);
                  if (kProd)
                    vv = xv;
                  else if (MODE == TM_RES)
                    vv = bq[q][r] - xv;
                  else
                    vv = xv + w * dq[q][r] * (bq[q][r] - xv);
                }
              }
              if (kAdd) vv += aq[q][r];
              if (MODE == TM_YDOT) dot += (double)xv * (double)vv;
              if (MODE == TM_JAC_ADD_DOT) dot += (double)vq[MODE == TM_JAC_ADD_DOT ? q : 0][r] * (double)vv;
              st32(y, d, r, vv);
            }
          }
        }
      }
      if (more) {
        store_u(ey + 2);
        store_e();
      }
      __syncthreads();
    }
  }
  if (kDot) {
    // fixed-order workgroup sum: waves, then wave partials in index order; the
    // wave partials reuse the (now idle) dynamic LDS, so the kernel's static LDS
    // stays zero and the full dynamic allowance remains available
    double* wsum = reinterpret_cast<double*>(lds_raw);
    __syncthreads();
    dot = wave_sum(dot);
    if (lane == 0) wsum[wave] = dot;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
#pragma unroll
      for (int v = 0; v < NT / 64; ++v) s += wsum[v];
      partial[blockIdx.x] = s;
    }
  }
}

static int g_wht_enable = 1;
static int g_wht_wgs = 1;
// tile per precision [float, double]: workgroup size and slices per layer of
// the sweeps (residual and Jacobi modes); y = A x products always take
// 16-slice tiles (fastest in both precisions)
static int g_wht_nt[2] = {1024, 512};
static int g_wht_ns[2] = {16, 8};
// per-host-thread overrides (0 = the process-wide value): lets concurrent
// solves on different threads use different tiles
static thread_local int t_wht_nt[2] = {0, 0};
static thread_local int t_wht_ns[2] = {0, 0};

static int wht_nt_eff(int p) { return t_wht_nt[p] ? t_wht_nt[p] : g_wht_nt[p]; }
static int wht_ns_eff(int p) { return t_wht_ns[p] ? t_wht_ns[p] : g_wht_ns[p]; }

void set_thread_wht_tiles(const int* nt, const int* ns) {
  for (int p = 0; p < 2; ++p) {
    t_wht_nt[p] = nt[p] == 0 ? 0 : ((nt[p] == 1024 || nt[p] == 256) ? nt[p] : 512);
    t_wht_ns[p] = ns[p] == 0 ? 0 : (ns[p] == 8 ? 8 : 16);
  }
}

static int wht_ns(size_t elem_bytes, int mode) {
  return (mode == TM_Y || mode == TM_YDOT) ? 16 : wht_ns_eff(elem_bytes == 8 ? 1 : 0);
}

void set_wht_threads(int is_double, int nt) {
  g_wht_nt[is_double ? 1 : 0] = (nt == 1024 || nt == 256) ? nt : 512;
}
int get_wht_threads(int is_double) { return g_wht_nt[is_double ? 1 : 0]; }
void set_wht_slices(int is_double, int ns) { g_wht_ns[is_double ? 1 : 0] = ns == 8 ? 8 : 16; }
int get_wht_slices(int is_double) { return g_wht_ns[is_double ? 1 : 0]; }

void set_wht(int on) { g_wht_enable = on ? 1 : 0; }
int get_wht() { return g_wht_enable; }
void set_wht_wgs(int n) { g_wht_wgs = n > 0 ? n : 1; }
int get_wht_wgs() { return g_wht_wgs; }

constexpr size_t kWMaxLds = 160 * 1024;

bool wht_supported(const Grid& g, size_t elem_bytes) {
  // 32-bit element indices and 32-bit byte offsets
  if (g.dim != 3 || g.ez > 128 || 3 * g.n_nodes() >= ((int64_t)1 << 31) ||
      3 * g.n_nodes() * (int64_t)elem_bytes >= ((int64_t)1 << 32) || g.n_elem() * (int64_t)elem_bytes >= ((int64_t)1 << 32))
    return false;
  for (const int ns : {16, wht_ns_eff(elem_bytes == 8 ? 1 : 0)}) {
    const int tc = wht_tc(g, ns);
    if (!((tc + 1) * g.nz() <= 96 * ns && tc * g.ez <= 64 * ns && (tc - 1) * g.nz() <= 64 * ns &&
          wht_lds_values(g, ns) * elem_bytes <= kWMaxLds))
      return false;
  }
  return true;
}

static int wht_device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

template <typename T, int MODE, int NT, int NS>
static void launch_wht_cfg(int n_wg, size_t lds, hipStream_t s, const MxArgs& A, const WhtK<T>& Gk, const T* E,
                           const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
                           double wscale, const T* xadd, T* y, double* partial, const T* dotv) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&op0_wht_kernel<T, MODE, NT, NS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kWMaxLds);
    attr_set = true;
  }
  op0_wht_kernel<T, MODE, NT, NS><<<n_wg, NT, lds, s>>>(Gk, A, E, cons, x, b, Dinv, omega, wscale, xadd, y,
                                                        partial, dotv);
}

// 16 slices: 1024 or 512 threads; 8 slices: 512 or 256 threads
template <typename T, int MODE>
static void launch_wht(int n_wg, size_t lds, hipStream_t s, const MxArgs& A, const WhtK<T>& Gk, const T* E,
                       const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
                       double wscale, const T* xadd, T* y, double* partial = nullptr,
                       const T* dotv = nullptr) {
  const int p = sizeof(T) == 8 ? 1 : 0;
  if (wht_ns(sizeof(T), MODE) == 16) {
    if (wht_nt_eff(p) == 1024)
      launch_wht_cfg<T, MODE, 1024, 16>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y, partial, dotv);
    else
      launch_wht_cfg<T, MODE, 512, 16>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y, partial, dotv);
  } else {
    if (wht_nt_eff(p) == 256)
      launch_wht_cfg<T, MODE, 256, 8>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y, partial, dotv);
    else
      launch_wht_cfg<T, MODE, 512, 8>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y, partial, dotv);
  }
}

// the table in the kernel's precision
template <typename T>
static WhtK<T> wht_k(const WhtTable& t) {
  WhtK<T> k;
  for (int i = 0; i < 48; ++i) k.g[i] = (T)t.g[i];
  return k;
}

// x tiles, persistent grid (ET_TUNE_WHT_WGS workgroups per compute unit for
// 16-slice tiles, twice as many for 8-slice tiles) and LDS bytes
static MxArgs wht_args(const Grid& g, int64_t e_sx, int64_t e_sy, size_t elem_bytes, int mode, int* n_wg,
                       size_t* lds) {
  MxArgs A;
  A.g = g;
  A.e_sx = e_sx;
  A.e_sy = e_sy;
  const int ns = wht_ns(elem_bytes, mode);
  const int tc = wht_tc(g, ns);
  A.n_xt = (g.nx() + tc - 2) / (tc - 1);
  A.nzg = 0;
  A.units = (int64_t)A.n_xt * g.ny();
  const int64_t slots = (int64_t)wht_device_cus() * g_wht_wgs * (16 / ns);
  *n_wg = (int)(A.units < slots ? A.units : slots);
  *lds = wht_lds_values(g, ns) * elem_bytes;
  return A;
}

int wht_partials(const Grid& g, size_t elem_bytes, int mode) {
  int n_wg = 0;
  size_t lds = 0;
  (void)wht_args(g, 0, 0, elem_bytes, mode, &n_wg, &lds);
  return n_wg;
}

template <typename T>
int op0_wht_ydot(const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E, const uint8_t* cons,
                 const T* x, T* y, double* partial, int max_partial, hipStream_t s) {
  ET_REQUIRE(wht_supported(g, sizeof(T)), "transformed-basis operator: 3-D grids with nelz <= 128 only");
  int n_wg = 0;
  size_t lds = 0;
  const MxArgs A = wht_args(g, e_sx, e_sy, sizeof(T), TM_YDOT, &n_wg, &lds);
  const WhtK<T> Gk = wht_k<T>(tab);
  ET_REQUIRE(partial != nullptr && n_wg <= max_partial, "transformed-basis operator: partial buffer too small");
  launch_wht<T, TM_YDOT>(n_wg, lds, s, A, Gk, E, cons, x, nullptr, nullptr, nullptr, 1.0, nullptr, y, partial);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int op0_wht_ydot<double>(const Grid&, int64_t, int64_t, const WhtTable&, const double*, const uint8_t*,
                                  const double*, double*, double*, int, hipStream_t);

template <typename T>
int op0_wht_jac_add_dot(const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E, const uint8_t* cons,
                        const T* x, const T* b, const T* Dinv, const double* omega, double wscale, const T* xadd,
                        T* y, const T* dotv, double* partial, int max_partial, hipStream_t s) {
  ET_REQUIRE(wht_supported(g, sizeof(T)), "transformed-basis operator: 3-D grids with nelz <= 128 only");
  int n_wg = 0;
  size_t lds = 0;
  const MxArgs A = wht_args(g, e_sx, e_sy, sizeof(T), TM_JAC_ADD_DOT, &n_wg, &lds);
  const WhtK<T> Gk = wht_k<T>(tab);
  ET_REQUIRE(partial != nullptr && n_wg <= max_partial, "transformed-basis operator: partial buffer too small");
  ET_REQUIRE(dotv != nullptr && xadd != nullptr, "transformed-basis operator: null operand");
  launch_wht<T, TM_JAC_ADD_DOT>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y, partial, dotv);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int op0_wht_jac_add_dot<float>(const Grid&, int64_t, int64_t, const WhtTable&, const float*, const uint8_t*,
                                        const float*, const float*, const float*, const double*, double,
                                        const float*, float*, const float*, double*, int, hipStream_t);
template int op0_wht_jac_add_dot<double>(const Grid&, int64_t, int64_t, const WhtTable&, const double*,
                                         const uint8_t*, const double*, const double*, const double*,
                                         const double*, double, const double*, double*, const double*, double*,
                                         int, hipStream_t);

template <typename T>
int op0_wht(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E,
            const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
            double wscale, const T* xadd, T* y, hipStream_t s) {
  ET_REQUIRE(wht_supported(g, sizeof(T)), "transformed-basis operator: 3-D grids with nelz <= 128 only");
  int n_wg = 0;
  size_t lds = 0;
  const MxArgs A = wht_args(g, e_sx, e_sy, sizeof(T), mode, &n_wg, &lds);
  const WhtK<T> Gk = wht_k<T>(tab);
  switch (mode) {
    case TM_Y: launch_wht<T, TM_Y>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_RES: launch_wht<T, TM_RES>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC: launch_wht<T, TM_JAC>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC_ADD: launch_wht<T, TM_JAC_ADD>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC0: launch_wht<T, TM_JAC0>(n_wg, lds, s, A, Gk, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    default: ET_REQUIRE(false, "transformed-basis operator: unknown mode");
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template int op0_wht<double>(int, const Grid&, int64_t, int64_t, const WhtTable&, const double*,
                             const uint8_t*, const double*, const double*, const double*,
                             const double*, double, const double*, double*, hipStream_t);
template int op0_wht<float>(int, const Grid&, int64_t, int64_t, const WhtTable&, const float*,
                            const uint8_t*, const float*, const float*, const float*,
                            const double*, double, const float*, float*, hipStream_t);

}  // namespace et
