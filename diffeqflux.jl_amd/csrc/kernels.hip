// Device kernels of the NeuralODE hot path for MI355X (gfx950 / CDNA4).
//
// * gemm_tn: one LDS-tiled MFMA GEMM template (BF16 v_mfma_f32_32x32x16_bf16
//   or FP32 v_mfma_f32_32x32x2f32, FP32 accumulation) with the Dense-layer
//   epilogues fused in: bias + activation (+ act', + transposed copy), the
//   Tsit5 stage combination and embedded-error partial sums, the adjoint
//   stage combination with the forward dense-output interpolation, and the
//   parameter-gradient quadrature fold.
// * single-workgroup controllers: FP64 PI step-size control, initial-dt
//   heuristic, tstop landing, saveat bookkeeping, retcodes.
// * elementwise kernels: solve initialisation, step begin (saves, FSAL,
//   cotangent jumps, mu commit), mu error norm, bias / split-K fold.
//
// Precision policies: FP32 (fp32 operands), BF16 (bf16 operands) and split
// BF16 (every operand stored as two bf16 planes hi + lo with hi = bf16(x),
// lo = bf16(x - hi); a GEMM runs three BF16 passes hi*hi + lo*hi + hi*lo over
// the same accumulators, which keeps ~16 mantissa bits of both operands).
//
// Layouts (all device buffers are allocated zero-filled; padded regions are
// never written with non-zeros):
//   state vectors      FP32 [Bp][D]            (Julia D x B column-major bytes)
//   GEMM activations   T    [planes][Bp][Fp]   (K contiguous for the next layer)
//   transposed copies  T    [planes][Fp][Bp]   (batch contiguous: K of weight grads)
//   weights            T    Wr[planes][out_p][in_p] (forward), Wc[planes][in_p][out_p]
// with Bp = B rounded up to 128 and Fp = F rounded up to 64.
#include <type_traits>

#include "node_internal.h"

namespace nodehip {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
// native 16-byte vector (HIP's uint4 is a struct wrapper that blocks SROA of
// the register-staged tiles)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__constant__ float cA[7][7];
__constant__ float cBT[7];

static const double kA[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365,
     -0.09249506636175525, 0, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159,
     -0.071584973281401, -0.028269050394068383, 0, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
     -3.290069515436081, 2.324710524099774, 0}};
static const double kBT[7] = {-0.00178001105222577714, -0.0008164344596567469,
                              0.007880878010261995,    -0.1447110071732629,
                              0.5823571654525552,      -0.45808210592918697,
                              0.015151515151515152};
__device__ const double dC[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
// free 4th-order dense output: b1(t) = t(r11 + t(r12 + t(r13 + t r14))),
// bi(t) = t^2 (ri2 + t(ri3 + t ri4))
__device__ const double dR1[4] = {1.0, -2.763706197274826, 2.9132554618219126,
                                  -1.0530884977290216};
__device__ const double dRI[6][3] = {
    {0.13169999999999998, -0.2234, 0.1017},
    {3.9302962368947516, -5.941033872131505, 2.490627285651253},
    {-12.411077166933676, 30.33818863028232, -16.548102889244902},
    {37.50931341651104, -88.1789048947664, 47.37952196281928},
    {-27.896526289197286, 65.09189467479366, -34.8706578614966},
    {1.5, -4.0, 2.5}};

// Stream-ordered on the library stream (a non-blocking stream does not wait
// for the null stream, and a copy from pageable memory may return before its
// DMA lands); waits for the copies before the host arrays go out of scope.
hipError_t upload_tableau(hipStream_t s) {
  float a[7][7], bt[7];
  for (int i = 0; i < 7; ++i) {
    for (int j = 0; j < 7; ++j) a[i][j] = (float)kA[i][j];
    bt[i] = (float)kBT[i];
  }
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(cA), a, sizeof(a), 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(cBT), bt, sizeof(bt), 0, hipMemcpyHostToDevice, s);
  const hipError_t w = hipStreamSynchronize(s);
  return e != hipSuccess ? e : w;
}

// PI controller constants (OrdinaryDiffEq defaults for Tsit5)
#define BETA1 0.14
#define BETA2 0.08
#define GAMMA 0.9
#define QMIN 0.2
#define QMAX 10.0
#define QOLDINIT 1e-4

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ bool gate_ok(const Ctl* c, int gate) {
  if (gate == GATE_ALWAYS) return true;
  if (c->done) return false;
  if (gate == GATE_FSAL) return c->fsal_reeval != 0;
  return true;
}

// Tape entry of forward step n: the entries form a ring of c->tape_ring slots
// when that is set (a forward solve that keeps only checkpoints: the stage
// kernels touch entries n and n + 1, the pending saves n - 1 and n), else the
// record of every step.
__device__ __forceinline__ int64_t tape_slot(const Ctl* c, int64_t n) {
  const int64_t r = c->tape_ring;
  return r > 0 ? n % r : n;
}

// Activations from the hardware exp2 / log2 / reciprocal (v_exp_f32,
// v_log_f32, v_rcp_f32, ~1 ulp each): the epilogues run them once per output
// element at low occupancy, where the full-range library forms dominated the
// kernel time.  Absolute error stays at the FP32 rounding level of the result.
#define LOG2E_F 1.4426950408889634f
#define LN2_F 0.6931471805599453f
__device__ __forceinline__ float fast_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z * LOG2E_F));
}
__device__ __forceinline__ float act_f(int a, float z) {
  switch (a) {
    case 1: {  // tanh(z) = sign(z) (1 - t) / (1 + t), t = exp(-2|z|)
      const float t = __builtin_amdgcn_exp2f(-2.f * LOG2E_F * fabsf(z));
      return copysignf((1.f - t) * __builtin_amdgcn_rcpf(1.f + t), z);
    }
    case 2:  // softplus: log1p(exp(-|z|)) + max(z, 0)
      return fmaxf(z, 0.f) +
             LN2_F * __builtin_amdgcn_logf(1.f + __builtin_amdgcn_exp2f(-fabsf(z) * LOG2E_F));
    case 3: return fmaxf(z, 0.f);
    case 4: return fast_sigmoid(z);
    default: return z;
  }
}
__device__ __forceinline__ float act_d(int a, float z, float y) {
  switch (a) {
    case 1: return 1.f - y * y;
    case 2: return fast_sigmoid(z);
    case 3: return z > 0.f ? 1.f : 0.f;
    case 4: return y * (1.f - y);
    default: return 1.f;
  }
}
__device__ __forceinline__ float map_f(int m, float u) { return m == 1 ? u * u * u : u; }

template <typename T> __device__ __forceinline__ T cvt(float x);
template <> __device__ __forceinline__ float cvt<float>(float x) { return x; }
template <> __device__ __forceinline__ bf16 cvt<bf16>(float x) { return (bf16)x; }
template <typename T> __device__ __forceinline__ float tof(T x);
template <> __device__ __forceinline__ float tof<float>(float x) { return x; }
template <> __device__ __forceinline__ float tof<bf16>(bf16 x) { return (float)x; }

// store 4 consecutive values (8 B for bf16, 16 B for fp32; caller guarantees alignment)
__device__ __forceinline__ void store4(float* p, const float v[4]) {
  *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void store4(bf16* p, const float v[4]) {
  typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
  bf16x4 o = {(bf16)v[0], (bf16)v[1], (bf16)v[2], (bf16)v[3]};
  *(bf16x4*)p = o;
}

// precision policy: operand type, split (hi/lo planes), act' storage type
template <typename T_, bool SPLIT_>
struct Prec {
  using T = T_;
  static constexpr bool SPLIT = SPLIT_;
  using ST = typename std::conditional<SPLIT_, float, T_>::type;
};
using PF32 = Prec<float, false>;
using PBF16 = Prec<bf16, false>;
using PBF16X3 = Prec<bf16, true>;

template <class PP>
__device__ __forceinline__ void put(typename PP::T* p, int64_t plane, float v) {
  if constexpr (PP::SPLIT) {
    const bf16 h = (bf16)v;
    p[0] = h;
    p[plane] = (bf16)(v - (float)h);
  } else {
    p[0] = cvt<typename PP::T>(v);
  }
}
template <class PP>
__device__ __forceinline__ float get(const typename PP::T* p, int64_t plane) {
  if constexpr (PP::SPLIT) return (float)p[0] + (float)p[plane];
  else return tof<typename PP::T>(p[0]);
}
template <class PP>
__device__ __forceinline__ void put4(typename PP::T* p, int64_t plane, const float v[4]) {
  if constexpr (PP::SPLIT) {
    float hi[4], lo[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      hi[i] = (float)(bf16)v[i];
      lo[i] = v[i] - hi[i];
    }
    store4(p, hi);
    store4(p + plane, lo);
  } else {
    store4(p, v);
  }
}

#define DISPATCH_PREC(prec, ...)                   \
  do {                                             \
    if ((prec) == 1) { using PP = PBF16; __VA_ARGS__; }   \
    else if ((prec) == 2) { using PP = PBF16X3; __VA_ARGS__; } \
    else { using PP = PF32; __VA_ARGS__; }          \
  } while (0)

// Block-wide reduction of (sum, max-flag); thread 0 writes part[2*slot..].
__device__ void block_reduce_store(double s, double f, double* part, int slot) {
  __shared__ double red[2][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    f = fmax(f, __shfl_xor(f, off));
  }
  if (lane == 0) { red[0][wave] = s; red[1][wave] = f; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0, ff = 0.0;
    for (int w = 0; w < nw; ++w) { ss += red[0][w]; ff = fmax(ff, red[1][w]); }
    part[2 * slot] = ss;
    part[2 * slot + 1] = ff;
  }
}

__device__ __forceinline__ double dspacing(double t) {
  double a = fabs(t);
  return nextafter(a, INFINITY) - a;
}

// ---------------------------------------------------------------- MFMA
template <typename T> struct Mma;
template <> struct Mma<bf16> {
  static __device__ __forceinline__ void run(f32x16& c, const u32x4& a, const u32x4& b) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
};
// FP32: 4 MFMA 32x32x2 per 16-B fragment.  Lane l supplies k = 4*(l/32) + j to
// step j for both operands, so the four steps cover the 8 k of the chunk.
template <> struct Mma<float> {
  static __device__ __forceinline__ void run(f32x16& c, const u32x4& a, const u32x4& b) {
    const f32x4v fa = __builtin_bit_cast(f32x4v, a), fb = __builtin_bit_cast(f32x4v, b);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(fa.x, fb.x, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(fa.y, fb.y, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(fa.z, fb.z, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(fa.w, fb.w, c, 0, 0, 0);
  }
};

// Accumulator layout of a 32x32 MFMA tile: lane l, register v holds
// row 8*(v/4) + 4*(l/32) + v%4, column l%32.
#define EPI_LOOP_BEGIN                                                     \
  _Pragma("unroll") for (int mi = 0; mi < MR; ++mi)                       \
  _Pragma("unroll") for (int mj = 0; mj < MC; ++mj) {                     \
    const int c = cb + mj * 32 + (lane & 31);                              \
    _Pragma("unroll") for (int q = 0; q < 4; ++q) {                        \
      const int r0 = rb + mi * 32 + q * 8 + (lane >> 5) * 4;
#define EPI_LOOP_END }}

// ---------------------------------------------------------------- tile staging
// Where a workgroup's tile sits and which wave block this thread owns.
struct TileGeo {
  int slot;         // tile index (partial-sum slot of the stage epilogues)
  int row0, col0;   // tile origin
  int rb, cb;       // this wave's block origin (absolute)
  int lr0, lc0;     // this wave's block origin inside the tile
};

typedef bf16 bf16x8v __attribute__((ext_vector_type(8)));

// Copy a rows x cols FP32 tile staged in LDS (row pitch `pitch` floats) to
// global memory as OT (16-byte stores along cols); SPLIT writes the bf16
// rounding residual to the second plane (dst + plane).
template <typename OT, bool SPLIT, int NT>
__device__ __forceinline__ void tile_copy_out(const float* lt, int pitch, int rows, int cols,
                                              OT* dst, int64_t ld, int64_t plane) {
  constexpr int E = 16 / sizeof(OT);
  const int cpr = cols / E;
  for (int i = threadIdx.x; i < rows * cpr; i += NT) {
    const int r = i / cpr, cc = (i - r * cpr) * E;
    const float* src = lt + r * pitch + cc;
    OT* d = dst + (int64_t)r * ld + cc;
    if constexpr (sizeof(OT) == 4) {
      *(f32x4v*)d = *(const f32x4v*)src;
    } else {
      const f32x4v a = *(const f32x4v*)src, b = *(const f32x4v*)(src + 4);
      const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
      bf16x8v hi, lo;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        hi[k] = (bf16)v[k];
        if constexpr (SPLIT) lo[k] = (bf16)(v[k] - (float)hi[k]);
      }
      *(bf16x8v*)d = hi;
      if constexpr (SPLIT) *(bf16x8v*)(d + plane) = lo;
    }
  }
}

// Stage this thread's accumulator-layout values (vals[mi][mj][v]) into LDS as
// a float tile, row-major [r][c] (TRANS = false) or transposed [c][r].
template <int MR, int MC, bool TRANS>
__device__ __forceinline__ void tile_stage(float* lt, int pitch, const f32x16 (&vals)[MR][MC],
                                           const TileGeo& g, int lane) {
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.lc0 + mj * 32 + (lane & 31);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = g.lr0 + mi * 32 + q * 8 + (lane >> 5) * 4;
        if constexpr (TRANS) {
          *(f32x4v*)(lt + c * pitch + r0) =
              f32x4v{vals[mi][mj][q * 4], vals[mi][mj][q * 4 + 1], vals[mi][mj][q * 4 + 2],
                     vals[mi][mj][q * 4 + 3]};
        } else {
#pragma unroll
          for (int p = 0; p < 4; ++p) lt[(r0 + p) * pitch + c] = vals[mi][mj][q * 4 + p];
        }
      }
    }
}

// Write a BR x BC tile of values to a row-major [r][c] buffer (ld) and/or its
// transposed [c][r] buffer (ldt) through LDS.  All threads must call it.
template <typename OT, bool SPLIT, int BR, int BC, int NT, int MR, int MC>
__device__ void tile_write(unsigned char* lds, const f32x16 (&vals)[MR][MC], const TileGeo& g,
                           int lane, OT* rowmajor, int64_t ld, int64_t plane, OT* trans,
                           int64_t ldt, int64_t tplane) {
  float* lt = (float*)lds;
  if (rowmajor) {
    __syncthreads();
    tile_stage<MR, MC, false>(lt, BC + 4, vals, g, lane);
    __syncthreads();
    tile_copy_out<OT, SPLIT, NT>(lt, BC + 4, BR, BC, rowmajor + (int64_t)g.row0 * ld + g.col0, ld,
                                 plane);
  }
  if (trans) {
    __syncthreads();
    tile_stage<MR, MC, true>(lt, BR + 4, vals, g, lane);
    __syncthreads();
    tile_copy_out<OT, SPLIT, NT>(lt, BR + 4, BC, BR, trans + (int64_t)g.col0 * ldt + g.row0, ldt,
                                 tplane);
  }
}

// ---------------------------------------------------------------- epilogues
template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiHidden& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& g, int lane, const Ctl*, int,
                                          unsigned char* lds) {
  using T = typename PP::T;
  using ST = typename PP::ST;
  const bool want_d = e.S != nullptr;
  f32x16 sv[MR][MC];
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.cb + mj * 32 + (lane & 31);
      const bool cv = c < e.C;
      const float bias = cv ? e.bias[c] : 0.f;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const float z = acc[mi][mj][v] + bias;
        const float y = act_f(e.act, z);
        sv[mi][mj][v] = (cv && want_d) ? act_d(e.act, z, y) : 0.f;
        acc[mi][mj][v] = cv ? y : 0.f;
      }
    }
  tile_write<T, PP::SPLIT, BR, BC, NT>(lds, acc, g, lane, (T*)e.Y, e.ldo, e.py, (T*)e.YT, e.ldt,
                                       e.pyt);
  if (want_d)
    tile_write<ST, false, BR, BC, NT>(lds, sv, g, lane, (ST*)e.S, e.ldo, 0, (ST*)nullptr, 0, 0);
}

// act' from the activation's output y (identity, softplus, relu):
// softplus'(z) = sigmoid(z) = 1 - exp(-y) with y = softplus(z) > 0; below
// y = 1/4 its Taylor series to y^6 (relative truncation < 5e-8, no
// cancellation), above it the hardware exp2 (relative error < 1e-6)
__device__ __forceinline__ float act_d_y(int a, float y) {
  switch (a) {
    case 2: {
      if (y < 0.25f) {
        float t = 1.f - y * (1.f / 6.f);
        t = 1.f - y * (1.f / 5.f) * t;
        t = 1.f - y * 0.25f * t;
        t = 1.f - y * (1.f / 3.f) * t;
        t = 1.f - y * 0.5f * t;
        return y * t;
      }
      return 1.f - __builtin_amdgcn_exp2f(-y * LOG2E_F);
    }
    case 3: return y > 0.f ? 1.f : 0.f;
    default: return 1.f;
  }
}

// activations whose derivative the epilogues derive from the output
__device__ __forceinline__ bool act_from_output_d(int a) { return a == 0 || a == 2 || a == 3; }

// act' of this thread's accumulator
// elements, loaded while the GEMM's last K tile is computed (padded columns
// read the zero-filled padding)
template <class PP, int MR, int MC>
__device__ __forceinline__ void bwd_data_prefetch(const EpiBwdData& e, f32x16 (&pre)[MR][MC],
                                                  const TileGeo& g, int lane) {
  using ST = typename PP::ST;
  const ST* S = (const ST*)e.S;
  // (the output-derived form is loaded by the epilogue itself: its two
  // planes would not fit the main loop's registers)
  if (e.yact >= 0) return;
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.cb + mj * 32 + (lane & 31);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = g.rb + mi * 32 + q * 8 + (lane >> 5) * 4;
#pragma unroll
        for (int p = 0; p < 4; ++p)
          pre[mi][mj][q * 4 + p] = tof<ST>(S[(int64_t)(r0 + p) * e.ldo + c]);
      }
    }
}

// Output-derived act': the tile's Y planes copied global -> LDS by the
// direct-to-LDS loads (4 bytes per lane, a dense [plane][BR][BC] tile) into
// the LDS buffer that the GEMM's last K tile does not use.
template <class PP, int BR, int BC>
constexpr int ylds_bytes() { return (PP::SPLIT ? 2 : 1) * BR * BC * (int)sizeof(typename PP::T); }

template <class PP, int BR, int BC, int NT>
__device__ __forceinline__ void bwd_data_prefetch_lds(const EpiBwdData& e, unsigned char* dst,
                                                      const TileGeo& g) {
  using T = typename PP::T;
  constexpr int NPL = PP::SPLIT ? 2 : 1;
  constexpr int EPL = 4 / (int)sizeof(T);                 // elements per lane
  constexpr int NI = BR * BC * (int)sizeof(T) / 256;      // wave loads per plane
  static_assert(BR * BC * sizeof(T) % 256 == 0, "tile / wave-load mismatch");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const T* Y = (const T*)e.Y;
  for (int i = wave; i < NI * NPL; i += NT / 64) {
    const int p = i / NI, ii = i - p * NI;
    const int el = (ii * 64 + lane) * EPL;
    const int r = el / BC, c = el - r * BC;
    const T* src = Y + p * e.py + (int64_t)(g.row0 + r) * e.ldo + g.col0 + c;
    __builtin_amdgcn_global_load_lds(
        (const void*)src,
        (__attribute__((address_space(3))) void*)(dst + p * (BR * BC * (int)sizeof(T)) + ii * 256),
        4, 0, 0);
  }
}

// optional bias column sums, then the row-major + transposed delta tiles
template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void bwd_data_finish(const EpiBwdData& e, f32x16 (&acc)[MR][MC],
                                                const TileGeo& g, int lane, unsigned char* lds) {
  using T = typename PP::T;
  if (e.bsum) {
    // column sums of the tile (fixed order: registers, lane halves, waves)
    constexpr int WR = BR / (MR * 32);
    float* red = (float*)lds;  // [WR][BC]
    const int wr = g.lr0 / (BR / WR);
    __syncthreads();
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      float cs = 0.f;
#pragma unroll
      for (int mi = 0; mi < MR; ++mi)
#pragma unroll
        for (int v = 0; v < 16; ++v) cs += acc[mi][mj][v];
      cs += __shfl_xor(cs, 32);
      if (lane < 32) red[wr * BC + g.lc0 + mj * 32 + lane] = cs;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < BC; i += NT) {
      float sm = 0.f;
#pragma unroll
      for (int w = 0; w < WR; ++w) sm += red[w * BC + i];
      e.bsum[(int64_t)(g.row0 / BR) * e.ldo + g.col0 + i] = sm;
    }
  }
  tile_write<T, PP::SPLIT, BR, BC, NT>(lds, acc, g, lane, (T*)e.Dl, e.ldo, e.pd, (T*)e.DlT,
                                       e.ldt, e.pdt);
}

template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiBwdData& e, f32x16 (&acc)[MR][MC],
                                          const f32x16 (&pre)[MR][MC], const TileGeo& g,
                                          int lane, unsigned char* lds, const unsigned char* ylds) {
  using T = typename PP::T;
  if (e.yact >= 0 && ylds) __syncthreads();  // direct loads done (vmcnt) and visible
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.cb + mj * 32 + (lane & 31);
      const bool cv = c < e.C;
      if (e.yact >= 0 && ylds) {
        const T* yt = (const T*)ylds;
        constexpr int PL = BR * BC;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r0 = g.lr0 + mi * 32 + q * 8 + (lane >> 5) * 4;
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            const T* ye = yt + (r0 + p) * BC + (c - g.col0);
            const float y = PP::SPLIT ? tof<T>(ye[0]) + tof<T>(ye[PL]) : tof<T>(ye[0]);
            acc[mi][mj][q * 4 + p] = cv ? acc[mi][mj][q * 4 + p] * act_d_y(e.yact, y) : 0.f;
          }
        }
      } else if (e.yact >= 0) {
        const T* Y = (const T*)e.Y;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r0 = g.rb + mi * 32 + q * 8 + (lane >> 5) * 4;
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            const float y = get<PP>(Y + (int64_t)(r0 + p) * e.ldo + c, e.py);
            acc[mi][mj][q * 4 + p] = cv ? acc[mi][mj][q * 4 + p] * act_d_y(e.yact, y) : 0.f;
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[mi][mj][v] = cv ? acc[mi][mj][v] * pre[mi][mj][v] : 0.f;
      }
    }
  bwd_data_finish<PP, BR, BC, NT>(e, acc, g, lane, lds);
}

// FFJORD adjoint: primal data gradient plus the second-order source (act'
// loaded in the epilogue; the source read per element)
template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiBwdSrc& es, f32x16 (&acc)[MR][MC],
                                          const TileGeo& g, int lane, const Ctl*, int,
                                          unsigned char* lds) {
  using T = typename PP::T;
  using ST = typename PP::ST;
  const EpiBwdData& e = es.d;
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.cb + mj * 32 + (lane & 31);
      const bool cv = c < e.C;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = g.rb + mi * 32 + q * 8 + (lane >> 5) * 4;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int64_t o = (int64_t)(r0 + p) * e.ldo + c;
          const float d = e.yact >= 0 ? act_d_y(e.yact, get<PP>((const T*)e.Y + o, e.py))
                                      : tof<ST>(((const ST*)e.S)[o]);
          const int v = q * 4 + p;
          acc[mi][mj][v] = cv ? acc[mi][mj][v] * d + es.src[o] : 0.f;
        }
      }
    }
  bwd_data_finish<PP, BR, BC, NT>(e, acc, g, lane, lds);
}

// FFJORD adjoint: tangent-backward chain (see EpiTanBwd)
template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiTanBwd& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& g, int lane, const Ctl*, int,
                                          unsigned char* lds) {
  using T = typename PP::T;
  using ST = typename PP::ST;
  const int yact = act_from_output_d(e.act) ? e.act : -1;
#pragma unroll
  for (int mi = 0; mi < MR; ++mi)
#pragma unroll
    for (int mj = 0; mj < MC; ++mj) {
      const int c = g.cb + mj * 32 + (lane & 31);
      const bool cv = c < e.C;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = g.rb + mi * 32 + q * 8 + (lane >> 5) * 4;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int64_t o = (int64_t)(r0 + p) * e.ldo + c;
          const float y = get<PP>((const T*)e.Y + o, e.py);
          const float d = yact >= 0 ? act_d_y(yact, y) : tof<ST>(((const ST*)e.S)[o]);
          const float t = get<PP>((const T*)e.Tf + o, e.ptf);
          float sq = 0.f;  // adot * act''
          if (e.act == 1) sq = -2.f * y * t;
          else if (e.act == 2) sq = t * (1.f - d);
          else if (e.act == 4) sq = t * (1.f - 2.f * y);
          const int v = q * 4 + p;
          const float tb = acc[mi][mj][v];
          e.src[o] = (e.src_acc ? e.src[o] : 0.f) + (cv ? tb * sq : 0.f);
          acc[mi][mj][v] = cv ? tb * d : 0.f;
        }
      }
    }
  tile_write<T, PP::SPLIT, BR, BC, NT>(lds, acc, g, lane, (T*)e.Ab, e.ldo, e.pab, (T*)e.AbT,
                                       e.ldt, e.pabt);
}

template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiFwdStage& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& geo, int lane, const Ctl* ctl,
                                          int , unsigned char*) {
  const int rb = geo.rb, cb = geo.cb;
  using T = typename PP::T;
  const int64_t n = (e.mode == SM_STEP) ? ctl->n : 0;
  float* ent = e.tape + tape_slot(ctl, n) * e.ent;
  float* nxt = e.tape + tape_slot(ctl, n + 1) * e.ent;  // entry n + 1
  const float dt = (float)ctl->dt;
  const float atol = (float)ctl->abstol, rtol = (float)ctl->reltol;
  const int s = e.stage;
  T* X0 = (T*)e.X0;
  double sum = 0.0, flag = 0.0;
  if (e.mode == SM_STEP) {
    // pass 1: loads and arithmetic only (no stores between the loads, which
    // the compiler could not reorder across: the stage slots share one array)
    f32x16 uv[MR][MC];
    EPI_LOOP_BEGIN
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        uv[mi][mj][v] = 0.f;
        if (c < e.D) {
          const int64_t idx = (int64_t)r * e.D + c;
          const float k = act_f(e.act, acc[mi][mj][v] + e.bias[c]);
          acc[mi][mj][v] = k;
          if (s < 6) {
            float sm = cA[s + 1][s] * k;
#pragma unroll
            for (int j = 0; j < 5; ++j)
              if (j < s) sm += cA[s + 1][j] * ent[(1 + j) * e.sv + idx];
            uv[mi][mj][v] = ent[idx] + dt * sm;
          } else {
            float ut = cBT[6] * k;
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * ent[(1 + j) * e.sv + idx];
            ut *= dt;
            const float up = ent[idx], un = nxt[idx];
            const float sk = atol + fmaxf(fabsf(up), fabsf(un)) * rtol;
            if (r < e.B) {
              const double a = (double)(ut / sk);
              sum += a * a;
              if (isnan(un)) flag = 1.0;
            }
          }
        }
      }
    EPI_LOOP_END
    // pass 2: stores
    EPI_LOOP_BEGIN
      if (c < e.D) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int r = r0 + p, v = q * 4 + p;
          const int64_t idx = (int64_t)r * e.D + c;
          const float k = acc[mi][mj][v];
          ent[(1 + s) * e.sv + idx] = k;
          if (s < 6) {
            const float u = uv[mi][mj][v];
            put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, map_f(e.map, u));
            if (s == 5) nxt[idx] = u;            // candidate tape[n+1].uprev
          } else {
            nxt[e.sv + idx] = k;                  // FSAL: tape[n+1].k1
          }
        }
      }
    EPI_LOOP_END
    if (e.part) block_reduce_store(sum, flag, e.part, geo.slot);
    return;
  }
  EPI_LOOP_BEGIN
    if (c < e.D) {
      const float bias = e.bias[c];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p;
        const int64_t idx = (int64_t)r * e.D + c;
        const float k = act_f(e.act, acc[mi][mj][q * 4 + p] + bias);
        const bool live = r < e.B;
        if (e.mode == SM_STEP) {
          ent[(1 + s) * e.sv + idx] = k;
          if (s < 6) {
            float sm = cA[s + 1][s] * k;
#pragma unroll
            for (int j = 0; j < 5; ++j)
              if (j < s) sm += cA[s + 1][j] * ent[(1 + j) * e.sv + idx];
            const float u = ent[idx] + dt * sm;
            put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, map_f(e.map, u));
            if (s == 5) nxt[idx] = u;            // candidate tape[n+1].uprev
          } else {
            nxt[e.sv + idx] = k;                  // FSAL: tape[n+1].k1
            float ut = cBT[6] * k;
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * ent[(1 + j) * e.sv + idx];
            ut *= dt;
            const float up = ent[idx], un = nxt[idx];
            const float sk = atol + fmaxf(fabsf(up), fabsf(un)) * rtol;
            if (live) {
              const double a = (double)(ut / sk);
              sum += a * a;
              if (isnan(un)) flag = 1.0;
            }
          }
        } else if (e.mode == SM_INIT_F0) {
          ent[e.sv + idx] = k;
          const float u0 = ent[idx];
          const float sk = atol + fabsf(u0) * rtol;
          if (live) { const double a = (double)(k / sk); sum += a * a; }
        } else if (e.mode == SM_INIT_F1) {
          e.scratch[idx] = k;
          const float u0 = ent[idx], f0 = ent[e.sv + idx];
          const float sk = atol + fabsf(u0) * rtol;
          if (live) {
            const double a = (double)((k - f0) / sk);
            sum += a * a;
            if (k != f0) flag = 1.0;
          }
        } else {
          e.scratch[idx] = k;
        }
      }
    }
  EPI_LOOP_END
  if (e.part) block_reduce_store(sum, flag, e.part, geo.slot);
}

template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiBwdLast& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& geo, int lane, const Ctl* ctl,
                                          int , unsigned char* lds) {
  const int rb = geo.rb, cb = geo.cb;
  using T = typename PP::T;
  const float dt = (float)ctl->dt;
  // next-stage inputs (lambda_s and x = map(u(t_s))), written through LDS
  // as row-major + transposed 16-byte stores after the element loop
  const bool stage_out = e.mode == SM_STEP && e.stage < 6;
  f32x16 lsv[MR][MC], xv[MR][MC];
#pragma unroll
  for (int i = 0; i < MR; ++i)
#pragma unroll
    for (int j = 0; j < MC; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) lsv[i][j][v] = xv[i][j][v] = 0.f;
  const float atol = (float)ctl->abstol, rtol = (float)ctl->reltol;
  const int s = e.stage;
  T* L = (T*)e.L;
  T* LT = (T*)e.LT;
  T* X0 = (T*)e.X0;
  T* X0T = (T*)e.X0T;
  double sum = 0.0, flag = 0.0;
  // interpolation data of the next stage's time
  const int slot = s + 1;
  int tj = 0, ex = 1;
  float tdt = 0.f, w[7] = {0, 0, 0, 0, 0, 0, 0};
  if (e.mode == SM_STEP && s < 6) {
    tj = ctl->st_step[slot];
    ex = ctl->st_exact[slot];
    tdt = ctl->st_dt[slot];
#pragma unroll
    for (int i = 0; i < 7; ++i) w[i] = ctl->st_w[slot][i];
  }
  if (e.mode == SM_STEP) {
    // pass 1: loads and arithmetic; pass 2: stores (see the forward stage)
    f32x16 uv[MR][MC];
    EPI_LOOP_BEGIN
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        uv[mi][mj][v] = 0.f;
        if (c < e.D) {
          const int64_t idx = (int64_t)r * e.D + c;
          float g = acc[mi][mj][v];
          if (e.map == 1) { const float u = e.Ucur[idx]; g *= 3.f * u * u; }
          const float k = -g;
          acc[mi][mj][v] = k;
          if (s < 6) {
            float sm = cA[s + 1][s] * k;
#pragma unroll
            for (int j = 0; j < 5; ++j)
              if (j < s) sm += cA[s + 1][j] * e.kl[j * e.sv + idx];
            lsv[mi][mj][v] = e.lam[idx] + dt * sm;
            const float* te = e.tape + (int64_t)tj * e.ent;
            float u = te[idx];
            if (!ex) {
              float sm2 = 0.f;
#pragma unroll
              for (int i = 0; i < 7; ++i) sm2 += w[i] * te[(1 + i) * e.sv + idx];
              u += tdt * sm2;
            }
            uv[mi][mj][v] = u;
            xv[mi][mj][v] = map_f(e.map, u);
          } else {
            float ut = cBT[6] * k;
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * e.kl[j * e.sv + idx];
            ut *= dt;
            const float lp = e.lam[idx], ln = e.lam_new[idx];
            const float sk = atol + fmaxf(fabsf(lp), fabsf(ln)) * rtol;
            if (r < e.B) {
              const double a = (double)(ut / sk);
              sum += a * a;
              if (isnan(ln)) flag = 1.0;
            }
          }
        }
      }
    EPI_LOOP_END
    EPI_LOOP_BEGIN
      if (c < e.D) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int r = r0 + p, v = q * 4 + p;
          const int64_t idx = (int64_t)r * e.D + c;
          e.kl[s * e.sv + idx] = acc[mi][mj][v];
          if (s < 6) {
            if (s == 5) e.lam_new[idx] = lsv[mi][mj][v];
            e.Ucur[idx] = uv[mi][mj][v];
          }
        }
      }
    EPI_LOOP_END
    if (stage_out) {
      tile_write<T, PP::SPLIT, BR, BC, NT, MR, MC>(lds, lsv, geo, lane, L, e.ldl, e.pl, LT, e.ldlt,
                                                   e.plt);
      tile_write<T, PP::SPLIT, BR, BC, NT, MR, MC>(lds, xv, geo, lane, X0, e.ldx, e.px, X0T,
                                                   e.ldxt, e.pxt);
    }
    if (e.part) block_reduce_store(sum, flag, e.part, geo.slot);
    return;
  }
  EPI_LOOP_BEGIN
    if (c < e.D) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p;
        const int64_t idx = (int64_t)r * e.D + c;
        float v = acc[mi][mj][q * 4 + p];
        if (e.map == 1) { const float u = e.Ucur[idx]; v *= 3.f * u * u; }
        const float k = -v;
        const bool live = r < e.B;
        if (e.mode == SM_PLAIN) {
          e.plain_out[idx] = v;
        } else if (e.mode == SM_INIT_F0) {
          e.kl[idx] = k;
          const float sk = atol + fabsf(e.lam[idx]) * rtol;
          if (live) { const double a = (double)(k / sk); sum += a * a; }
        } else if (e.mode == SM_INIT_F1) {
          e.kl[7 * e.sv + idx] = k;
          const float f0 = e.kl[idx];
          const float sk = atol + fabsf(e.lam[idx]) * rtol;
          if (live) {
            const double a = (double)((k - f0) / sk);
            sum += a * a;
            if (k != f0) flag = 1.0;
          }
        } else {
          e.kl[s * e.sv + idx] = k;
          if (s < 6) {
            float sm = cA[s + 1][s] * k;
#pragma unroll
            for (int j = 0; j < 5; ++j)
              if (j < s) sm += cA[s + 1][j] * e.kl[j * e.sv + idx];
            const float ls = e.lam[idx] + dt * sm;
            if (s == 5) e.lam_new[idx] = ls;
            lsv[mi][mj][q * 4 + p] = ls;
            const float* te = e.tape + (int64_t)tj * e.ent;
            float u = te[idx];
            if (!ex) {
              float sm2 = 0.f;
#pragma unroll
              for (int i = 0; i < 7; ++i) sm2 += w[i] * te[(1 + i) * e.sv + idx];
              u += tdt * sm2;
            }
            xv[mi][mj][q * 4 + p] = map_f(e.map, u);
            e.Ucur[idx] = u;
          } else {
            float ut = cBT[6] * k;
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * e.kl[j * e.sv + idx];
            ut *= dt;
            const float lp = e.lam[idx], ln = e.lam_new[idx];
            const float sk = atol + fmaxf(fabsf(lp), fabsf(ln)) * rtol;
            if (live) {
              const double a = (double)(ut / sk);
              sum += a * a;
              if (isnan(ln)) flag = 1.0;
            }
          }
        }
      }
    }
  EPI_LOOP_END
  if (stage_out) {
    tile_write<T, PP::SPLIT, BR, BC, NT, MR, MC>(lds, lsv, geo, lane, L, e.ldl, e.pl, LT, e.ldlt,
                                                 e.plt);
    tile_write<T, PP::SPLIT, BR, BC, NT, MR, MC>(lds, xv, geo, lane, X0, e.ldx, e.px, X0T, e.ldxt,
                                                 e.pxt);
  }
  if (e.part) block_reduce_store(sum, flag, e.part, geo.slot);
}

template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiCnfStage& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& geo, int lane, const Ctl* ctl,
                                          int , unsigned char* lds) {
  const int rb = geo.rb, cb = geo.cb;
  using T = typename PP::T;
  const int64_t n = (e.mode == SM_STEP) ? ctl->n : 0;
  float* ent = e.tape + tape_slot(ctl, n) * e.ent;
  float* nxt = e.tape + tape_slot(ctl, n + 1) * e.ent;  // entry n + 1
  const float dt = (float)ctl->dt;
  const float atol = (float)ctl->abstol, rtol = (float)ctl->reltol;
  const int s = e.stage;
  T* X0 = (T*)e.X0;
  // The GEMM gives w = e^T df/dx per sample (x the model input, w's first dm
  // columns the z part).  Per-row sums go through LDS (one tile spans every
  // column of a row: the plan keeps BC >= the width), one sum at a time.
  float* pl = (float*)lds;
  float* tr = pl + BR * (BC + 1);   // -> trace per row
  float* q1 = tr + BR;              // -> |f|^2 per row (regularize)
  float* q2 = q1 + BR;              // -> |w|^2 per row (regularize)
  const bool last = e.pass + 1 >= e.npass;
  if (e.npass > 1 && !last) {
    // exact trace, not the last unit probe: accumulate J[pass][pass] and
    // e[pass] * J[pass][:] (the regulariser's e^T J); no stage
    EPI_LOOP_BEGIN
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        if (c < e.dm) {
          const float w = acc[mi][mj][v];
          if (c == e.pass) e.trc[r] += w;
          if (e.reg) e.wacc[(int64_t)r * e.dm + c] += e.e[(int64_t)r * e.dm + e.pass] * w;
        }
      }
    EPI_LOOP_END
    return;
  }
  __syncthreads();
  // w_eff: this pass's w (Hutchinson), or the accumulated e^T J (exact)
  EPI_LOOP_BEGIN
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = r0 + p, v = q * 4 + p;
      float w = c < e.dm ? acc[mi][mj][v] : 0.f;
      if (e.npass > 1 && c < e.dm) {
        if (c == e.pass) { tr[r - geo.row0] = e.trc[r] + w; e.trc[r] = 0.f; }
        if (e.reg) {
          const int64_t o = (int64_t)r * e.dm + c;
          w = e.wacc[o] + e.e[(int64_t)r * e.dm + e.pass] * w;
          e.wacc[o] = 0.f;
        }
      }
      acc[mi][mj][v] = w;
      if (e.npass == 1)
        pl[(r - geo.row0) * (BC + 1) + (c - geo.col0)] = c < e.dm ? w * e.e[(int64_t)r * e.dm + c] : 0.f;
    }
  EPI_LOOP_END
  __syncthreads();
  if (e.npass == 1) {
    for (int i = threadIdx.x; i < BR; i += NT) {
      float sm = 0.f;
      for (int j = 0; j < BC; ++j) sm += pl[i * (BC + 1) + j];
      tr[i] = sm;
    }
    __syncthreads();
  }
  if (e.reg) {
    EPI_LOOP_BEGIN
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        const float fv = c < e.dm ? e.f[(int64_t)r * e.dm + c] : 0.f;
        pl[(r - geo.row0) * (BC + 1) + (c - geo.col0)] = fv * fv;
      }
    EPI_LOOP_END
    __syncthreads();
    for (int i = threadIdx.x; i < BR; i += NT) {
      float sm = 0.f;
      for (int j = 0; j < BC; ++j) sm += pl[i * (BC + 1) + j];
      q1[i] = sm;
    }
    __syncthreads();
    EPI_LOOP_BEGIN
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        const float w = acc[mi][mj][v];
        pl[(r - geo.row0) * (BC + 1) + (c - geo.col0)] = w * w;
      }
    EPI_LOOP_END
    __syncthreads();
    for (int i = threadIdx.x; i < BR; i += NT) {
      float sm = 0.f;
      for (int j = 0; j < BC; ++j) sm += pl[i * (BC + 1) + j];
      q2[i] = sm;
    }
    __syncthreads();
  }
  // stage values k = [f; -trace] (pass 1: loads / arithmetic, pass 2: stores)
  double sum = 0.0, flag = 0.0;
  f32x16 uv[MR][MC];
  EPI_LOOP_BEGIN
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int r = r0 + p, v = q * 4 + p;
      uv[mi][mj][v] = 0.f;
      acc[mi][mj][v] = 0.f;
      if (c < e.D) {
        const int64_t idx = (int64_t)r * e.D + c;
        const int rl = r - geo.row0;
        const float k = c < e.dm ? e.f[(int64_t)r * e.dm + c]
                        : c == e.dm ? -tr[rl] : c == e.dm + 1 ? q1[rl] : sqrtf(q2[rl]);
        acc[mi][mj][v] = k;
        const bool live = r < e.B;
        if (e.mode == SM_STEP) {
          if (s < 6) {
            float sm = cA[s + 1][s] * k;
#pragma unroll
            for (int j = 0; j < 5; ++j)
              if (j < s) sm += cA[s + 1][j] * ent[(1 + j) * e.sv + idx];
            uv[mi][mj][v] = ent[idx] + dt * sm;
          } else {
            float ut = cBT[6] * k;
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * ent[(1 + j) * e.sv + idx];
            ut *= dt;
            const float up = ent[idx], un = nxt[idx];
            const float sk = atol + fmaxf(fabsf(up), fabsf(un)) * rtol;
            if (live) {
              const double a = (double)(ut / sk);
              sum += a * a;
              if (isnan(un)) flag = 1.0;
            }
          }
        } else if (e.mode == SM_INIT_F0) {
          const float sk = atol + fabsf(ent[idx]) * rtol;
          if (live) { const double a = (double)(k / sk); sum += a * a; }
        } else if (e.mode == SM_INIT_F1) {
          const float f0 = ent[e.sv + idx];
          const float sk = atol + fabsf(ent[idx]) * rtol;
          if (live) {
            const double a = (double)((k - f0) / sk);
            sum += a * a;
            if (k != f0) flag = 1.0;
          }
        }
      }
    }
  EPI_LOOP_END
  const double tnext = e.mode == SM_STEP && s < 6 ? ctl->t + dC[s + 1] * ctl->dt : 0.0;
  EPI_LOOP_BEGIN
    if (c < e.D) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + p, v = q * 4 + p;
        const int64_t idx = (int64_t)r * e.D + c;
        const float k = acc[mi][mj][v];
        if (e.mode == SM_STEP) {
          ent[(1 + s) * e.sv + idx] = k;
          if (s < 6) {
            const float u = uv[mi][mj][v];
            if (c < e.dm) put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, u);
            else if (e.tcat && c == e.dm && r < e.B)
              put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, (float)tnext);
            if (s == 5) nxt[idx] = u;            // candidate tape[n+1].uprev
          } else {
            nxt[e.sv + idx] = k;                  // FSAL: tape[n+1].k1
          }
        } else if (e.mode == SM_INIT_F0) {
          ent[e.sv + idx] = k;
        }
      }
    }
  EPI_LOOP_END
  if (e.part) block_reduce_store(sum, flag, e.part, geo.slot);
}

__device__ __forceinline__ void fold_one(const EpiWgrad& e, float* gcur, float* gnext,
                                         int64_t idx, float g) {
  if (e.flags & WG_STORE_CUR) gcur[idx] = g;
  if (e.flags & WG_STORE_PTR) e.gptr[idx] = g;
  if (e.flags & WG_INIT) {
    const float g1 = gcur[idx];
    e.accb[idx] = e.cb1 * g1 + e.cb * g;
    e.acce[idx] = e.ce1 * g1 + e.ce * g;
  } else {
    if (e.flags & WG_B) e.accb[idx] += e.cb * g;
    if (e.flags & WG_E) e.acce[idx] += e.ce * g;
  }
  if (e.flags & WG_STORE_NEXT) gnext[idx] = g;
}

template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiWgrad& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& geo, int lane, const Ctl* ctl,
                                          int split, unsigned char*) {
  const int rb = geo.rb, cb = geo.cb;
  const int gsel = ctl->gsel;
  float* gcur = gsel ? e.gf1 : e.gf0;
  float* gnext = gsel ? e.gf0 : e.gf1;
  float* wp = e.wpart ? e.wpart + (int64_t)split * e.in * e.out : nullptr;
  EPI_LOOP_BEGIN
    if (c < e.out) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int i = r0 + p;
        if (i < e.in) {
          const float g = -acc[mi][mj][q * 4 + p];
          if (wp) wp[(int64_t)i * e.out + c] = g;       // reduced by the fold kernel
          else fold_one(e, gcur, gnext, e.w0 + (int64_t)i * e.out + c, g);
        }
      }
    }
  EPI_LOOP_END
}

// split-K partial products of the skinny stage GEMMs: split s stores its FP32
// tile at ws[s * pstride + r * ld + c]; the stage kernels below sum the splits
// in split order and run the per-element stage work
template <class PP, int BR, int BC, int NT, int MR, int MC>
__device__ __forceinline__ void epi_apply(const EpiPartial& e, f32x16 (&acc)[MR][MC],
                                          const TileGeo& geo, int lane, const Ctl*, int split,
                                          unsigned char*) {
  const int rb = geo.rb, cb = geo.cb;
  float* w = e.ws + (int64_t)split * e.pstride;
  EPI_LOOP_BEGIN
#pragma unroll
    for (int p = 0; p < 4; ++p) w[(int64_t)(r0 + p) * e.ld + c] = acc[mi][mj][q * 4 + p];
  EPI_LOOP_END
}

// ---------------------------------------------------------------- GEMM
// C[r][c] = sum_k A[r][k] * B[c][k]; rows = batch (or input feature for
// weight gradients), cols = output feature.  WRxWC waves per workgroup, each
// wave a (BR/WR)x(BC/WC) block of 32x32 MFMA tiles.  K is staged through LDS
// in 128-byte rows (64 bf16 / 32 fp32 per row), double-buffered, one barrier
// per K tile, with an S-deep register prefetch ring: while tile t is computed
// from LDS, the loads of tiles t+1..t+S are in flight (the loop is unrolled by
// S so every ring index is static).  Split BF16 loads the hi and lo planes of
// both operands per K tile and issues hi*hi + lo*hi + hi*lo MFMAs.
// Split-K (ksplit > 1): workgroup group s covers K tiles [s*nk, (s+1)*nk).
// Tile order is XCD-aware: consecutive tiles of the row-major tile grid go to
// the same XCD so that its L2 holds their operands.
template <class PP, int BR, int BC, int WR, int WC, int S, typename Epi>
__global__ void __launch_bounds__(WR * WC * 64)
    gemm_tn(GemmArgs g, Epi epi, const Ctl* __restrict__ ctl, int gate) {
  using T = typename PP::T;
  if (!gate_ok(ctl, gate)) return;
  constexpr int NT = WR * WC * 64;
  constexpr int PITCH = 144;
  constexpr int BKB = 128;  // bytes of K per tile row
  constexpr int MR = BR / WR / 32, MC = BC / WC / 32;
  constexpr int RP = NT / 8;  // rows per load pass
  constexpr int NPL = PP::SPLIT ? 2 : 1;        // operand planes per K tile
  constexpr int A_LOADS = BR / RP, B_LOADS = BC / RP;
  constexpr int ROWS = (BR + BC) * NPL;          // LDS rows per buffer
  static_assert(BR % RP == 0 && BC % RP == 0, "tile/thread mismatch");
  __shared__ __attribute__((aligned(16))) unsigned char lds[2 * ROWS * PITCH];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WC, wc = wave % WC;
  const int nt = g.row_tiles * g.col_tiles;
  const int split = blockIdx.x / nt;
  int bid = blockIdx.x - split * nt;
  if ((nt & 7) == 0) bid = (bid & 7) * (nt >> 3) + (bid >> 3);
  const int ct = bid % g.col_tiles, rt = bid / g.col_tiles;
  const int row0 = rt * BR, col0 = ct * BC;

  const int64_t ldab = g.lda * (int64_t)sizeof(T), ldbb = g.ldb * (int64_t)sizeof(T);
  const int64_t pab = g.pa * (int64_t)sizeof(T), pbb = g.pb * (int64_t)sizeof(T);
  const int lr = tid >> 3, lc = tid & 7;
  const unsigned char* Ag = (const unsigned char*)g.A + (int64_t)(row0 + lr) * ldab + lc * 16;
  const unsigned char* Bg = (const unsigned char*)g.B + (int64_t)(col0 + lr) * ldbb + lc * 16;

  u32x4 ra[S][NPL][A_LOADS], rbv[S][NPL][B_LOADS];
  f32x16 acc[MR][MC];
#pragma unroll
  for (int i = 0; i < MR; ++i)
#pragma unroll
    for (int j = 0; j < MC; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  const int nk = (int)((int64_t)g.K * sizeof(T) / BKB) / g.ksplit;
  const int k0 = split * nk;
  // global -> register ring slot (static slot index)
#define GEMM_GLOAD(t, slot)                                                        \
  do {                                                                             \
    const int64_t kb_ = (int64_t)(k0 + (t)) * BKB;                                 \
    _Pragma("unroll") for (int p_ = 0; p_ < NPL; ++p_) {                           \
      _Pragma("unroll") for (int i_ = 0; i_ < A_LOADS; ++i_)                       \
        ra[slot][p_][i_] = *(const u32x4*)(Ag + p_ * pab + (int64_t)(i_ * RP) * ldab + kb_); \
      _Pragma("unroll") for (int i_ = 0; i_ < B_LOADS; ++i_)                       \
        rbv[slot][p_][i_] = *(const u32x4*)(Bg + p_ * pbb + (int64_t)(i_ * RP) * ldbb + kb_); \
    }                                                                              \
  } while (0)
  // register ring slot -> LDS buffer; rows: [A plane 0 | A plane 1 | B plane 0 | B plane 1]
#define GEMM_SSTORE(buf, slot)                                                     \
  do {                                                                             \
    unsigned char* base_ = lds + (buf) * ROWS * PITCH + lr * PITCH + lc * 16;      \
    _Pragma("unroll") for (int p_ = 0; p_ < NPL; ++p_) {                           \
      _Pragma("unroll") for (int i_ = 0; i_ < A_LOADS; ++i_)                       \
        *(u32x4*)(base_ + (p_ * BR + i_ * RP) * PITCH) = ra[slot][p_][i_];         \
      _Pragma("unroll") for (int i_ = 0; i_ < B_LOADS; ++i_)                       \
        *(u32x4*)(base_ + (NPL * BR + p_ * BC + i_ * RP) * PITCH) = rbv[slot][p_][i_]; \
    }                                                                              \
  } while (0)

  TileGeo geo;
  geo.slot = bid;
  geo.row0 = row0;
  geo.col0 = col0;
  geo.lr0 = wr * (BR / WR);
  geo.lc0 = wc * (BC / WC);
  geo.rb = row0 + geo.lr0;
  geo.cb = col0 + geo.lc0;
  constexpr bool kPre = std::is_same<Epi, EpiBwdData>::value;
  constexpr bool kYlds = kPre && ylds_bytes<PP, BR, BC>() <= ROWS * PITCH;
  f32x16 pre[kPre ? MR : 1][kPre ? MC : 1];
  const int arow = wr * (BR / WR) + (lane & 31);
  const int brow = NPL * BR + wc * (BC / WC) + (lane & 31);
  const int koff = (lane >> 5) * 16;
  // one K tile from LDS buffer `buf` through the MFMAs
#define GEMM_COMPUTE(buf)                                                          \
  do {                                                                             \
    const unsigned char* base_ = lds + (buf) * ROWS * PITCH;                       \
    _Pragma("unroll") for (int kk = 0; kk < 4; ++kk) {                             \
      u32x4 fa[NPL][MR], fb[NPL][MC];                                              \
      _Pragma("unroll") for (int p = 0; p < NPL; ++p) {                            \
        _Pragma("unroll") for (int mi = 0; mi < MR; ++mi)                          \
          fa[p][mi] = *(const u32x4*)(base_ + (arow + p * BR + mi * 32) * PITCH + kk * 32 + koff); \
        _Pragma("unroll") for (int mj = 0; mj < MC; ++mj)                          \
          fb[p][mj] = *(const u32x4*)(base_ + (brow + p * BC + mj * 32) * PITCH + kk * 32 + koff); \
      }                                                                            \
      _Pragma("unroll") for (int mi = 0; mi < MR; ++mi)                            \
        _Pragma("unroll") for (int mj = 0; mj < MC; ++mj) {                        \
          Mma<T>::run(acc[mi][mj], fa[0][mi], fb[0][mj]);                          \
          if constexpr (PP::SPLIT) {                                               \
            Mma<T>::run(acc[mi][mj], fa[NPL - 1][mi], fb[0][mj]);                  \
            Mma<T>::run(acc[mi][mj], fa[0][mi], fb[NPL - 1][mj]);                  \
          }                                                                        \
        }                                                                          \
    }                                                                              \
  } while (0)
  // the data-gradient epilogue's operands, loaded during the last K tile
#define GEMM_EPI_PREFETCH()                                                        \
  do {                                                                             \
    if constexpr (kPre) {                                                          \
      bwd_data_prefetch<PP>(epi, pre, geo, lane);                                  \
      if constexpr (kYlds)                                                         \
        if (epi.yact >= 0)                                                         \
          bwd_data_prefetch_lds<PP, BR, BC, NT>(epi, lds + (nk & 1) * ROWS * PITCH, geo); \
    }                                                                              \
  } while (0)
  if (nk % S == 0 && nk >= 2 * S) {
    // Branch-free schedule: every steady-state iteration refills its ring
    // slot, the last S iterations load nothing.  With no load behind a
    // branch the compiler's wait counts stay exact (a conditional refill
    // makes it wait for the loads just issued), so S - 1 K tiles of loads
    // stay in flight behind the MFMAs.
#pragma unroll
    for (int j = 0; j < S; ++j) GEMM_GLOAD(j, j);
    GEMM_SSTORE(0, 0);
    __syncthreads();
    for (int t0 = 0; t0 < nk - S; t0 += S) {
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const int t = t0 + j;
        GEMM_GLOAD(t + S, j);  // slot j held tile t, already in LDS
        GEMM_COMPUTE(t & 1);
        GEMM_SSTORE((t + 1) & 1, (j + 1) % S);
        __syncthreads();
      }
    }
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int t = nk - S + j;
      if (j == S - 1) GEMM_EPI_PREFETCH();
      GEMM_COMPUTE(t & 1);
      if (j + 1 < S) GEMM_SSTORE((t + 1) & 1, (j + 1) % S);
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int j = 0; j < S; ++j)
      if (j < nk) GEMM_GLOAD(j, j);
    GEMM_SSTORE(0, 0);
    __syncthreads();
    for (int t0 = 0; t0 < nk; t0 += S) {
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const int t = t0 + j;
        if (t < nk) {
          // slot j held tile t, already in LDS: refill it with tile t+S
          if (t + S < nk) GEMM_GLOAD(t + S, j);
          if (t == nk - 1) GEMM_EPI_PREFETCH();
          GEMM_COMPUTE(t & 1);
          if (t + 1 < nk) GEMM_SSTORE((t + 1) & 1, (j + 1) % S);
          __syncthreads();
        }
      }
    }
  }
#undef GEMM_COMPUTE
#undef GEMM_EPI_PREFETCH
#undef GEMM_GLOAD
#undef GEMM_SSTORE
  if (g.ws != nullptr && g.ksplit > 1) {
    // split-K serial fixup: partial tile out (coalesced, thread-minor)
    constexpr int TILE = MR * MC * 16 * NT;
    float* mine = g.ws + ((int64_t)split * nt + bid) * TILE;
#pragma unroll
    for (int mi = 0; mi < MR; ++mi)
#pragma unroll
      for (int mj = 0; mj < MC; ++mj)
#pragma unroll
        for (int v = 0; v < 16; ++v) mine[((mi * MC + mj) * 16 + v) * NT + tid] = acc[mi][mj][v];
    __threadfence();
    __syncthreads();
    __shared__ int is_last;
    if (tid == 0) is_last = atomicAdd(g.cnt + bid, 1) == g.ksplit - 1;
    __syncthreads();
    if (!is_last) return;
    __threadfence();
#pragma unroll
    for (int mi = 0; mi < MR; ++mi)
#pragma unroll
      for (int mj = 0; mj < MC; ++mj)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int64_t o = (int64_t)((mi * MC + mj) * 16 + v) * NT + tid;
          float sm = 0.f;
          for (int sp = 0; sp < g.ksplit; ++sp) sm += g.ws[((int64_t)sp * nt + bid) * TILE + o];
          acc[mi][mj][v] = sm;
        }
    if (tid == 0) g.cnt[bid] = 0;
  }
  if constexpr (kPre)
    epi_apply<PP, BR, BC, NT>(epi, acc, pre, geo, lane, lds,
                              kYlds ? lds + (nk & 1) * ROWS * PITCH : nullptr);
  else epi_apply<PP, BR, BC, NT>(epi, acc, geo, lane, ctl, split, lds);
}

// tile configurations: {BR, BC, WR, WC}
static const int kCfg[6][4] = {{128, 128, 2, 2}, {128, 64, 2, 2}, {64, 64, 2, 2}, {32, 64, 1, 2},
                               {128, 128, 4, 2}, {128, 128, 2, 4}};
int gemm_cfg_count() { return 6; }
int gemm_cfg_tile_rows(int cfg) { return kCfg[cfg][0]; }
int gemm_cfg_tile_cols(int cfg) { return kCfg[cfg][1]; }

// register prefetch depth per tile: deeper for small FP32 / BF16 tiles (less
// MFMA work per K tile to cover the load latency); 2 for split BF16, which
// keeps the split stage GEMMs' short K loops (4..16 tiles) on the branch-free
// schedule (measured ~1% faster on the headline than 3 or 4)
template <class PP, int BR, int BC> struct Depth {
  static constexpr int value = (BR * BC >= 128 * 128) ? (PP::SPLIT ? 2 : 3) : (PP::SPLIT ? 2 : 4);
};

template <class PP, typename Epi>
static void launch_t(int cfg, const GemmArgs& g, const Epi& e, const Ctl* ctl, int gate,
                     hipStream_t s) {
  const int nb = g.row_tiles * g.col_tiles * (g.ksplit > 0 ? g.ksplit : 1);
  GemmArgs ga = g;
  if (ga.ksplit < 1) ga.ksplit = 1;
  // the stage epilogues (per-element Tsit5 / adjoint stage work) are planned
  // with at most 64 output columns per tile; their 128x128 form is not built
  constexpr bool kStage = std::is_same<Epi, EpiFwdStage>::value ||
                         std::is_same<Epi, EpiBwdLast>::value || std::is_same<Epi, EpiCnfStage>::value ||
                         std::is_same<Epi, EpiPartial>::value;
  if (kStage && (cfg == 0 || cfg >= 4)) cfg = 1;
  switch (cfg) {
    case 0:
      if constexpr (!kStage)
        hipLaunchKernelGGL((gemm_tn<PP, 128, 128, 2, 2, Depth<PP, 128, 128>::value, Epi>), dim3(nb), dim3(256), 0, s, ga, e, ctl, gate);
      break;
    case 1: hipLaunchKernelGGL((gemm_tn<PP, 128, 64, 2, 2, Depth<PP, 128, 64>::value, Epi>), dim3(nb), dim3(256), 0, s, ga, e, ctl, gate); break;
    case 2: hipLaunchKernelGGL((gemm_tn<PP, 64, 64, 2, 2, Depth<PP, 64, 64>::value, Epi>), dim3(nb), dim3(256), 0, s, ga, e, ctl, gate); break;
    case 3: hipLaunchKernelGGL((gemm_tn<PP, 32, 64, 1, 2, Depth<PP, 32, 64>::value, Epi>), dim3(nb), dim3(128), 0, s, ga, e, ctl, gate); break;
    case 4:
      if constexpr (!kStage)
        hipLaunchKernelGGL((gemm_tn<PP, 128, 128, 4, 2, Depth<PP, 128, 128>::value, Epi>), dim3(nb), dim3(512), 0, s, ga, e, ctl, gate);
      break;
    default:
      if constexpr (!kStage)
        hipLaunchKernelGGL((gemm_tn<PP, 128, 128, 2, 4, Depth<PP, 128, 128>::value, Epi>), dim3(nb), dim3(512), 0, s, ga, e, ctl, gate);
      break;
  }
}
void launch_gemm_hidden(int prec, int cfg, const GemmArgs& g, const EpiHidden& e,
                        const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_fwdstage(int prec, int cfg, const GemmArgs& g, const EpiFwdStage& e,
                          const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_bwddata(int prec, int cfg, const GemmArgs& g, const EpiBwdData& e,
                         const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_cnfstage(int prec, int cfg, const GemmArgs& g, const EpiCnfStage& e,
                          const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_bwdlast(int prec, int cfg, const GemmArgs& g, const EpiBwdLast& e,
                         const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_partial(int prec, int cfg, const GemmArgs& g, const EpiPartial& e,
                         const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_wgrad(int prec, int cfg, const GemmArgs& g, const EpiWgrad& e,
                       const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_tanbwd(int prec, int cfg, const GemmArgs& g, const EpiTanBwd& e,
                        const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }
void launch_gemm_bwdsrc(int prec, int cfg, const GemmArgs& g, const EpiBwdSrc& e,
                        const Ctl* ctl, int gate, hipStream_t s) { DISPATCH_PREC(prec, launch_t<PP>(cfg, g, e, ctl, gate, s)); }

// ---------------------------------------------------------------- stage kernels
// Per-element stage work of the skinny last-layer GEMMs, run over the summed
// split-K partials by the whole chip: a ST_THREADS block takes tiles of
// ST_ROWS rows x 64 columns (thread: column tid % 64, rows RPT (tid / 64) + j;
// transposed outputs: column tid / ST_G, rows RPT (tid % ST_G) + j, so each
// column's rows of a tile are one contiguous run).  The arithmetic per element
// is that of the fused epilogues (EpiFwdStage, EpiBwdLast); a thread issues
// the loads of all its rows before any store.
constexpr int ST_THREADS = 1024;
constexpr int ST_ROWS = 16;
constexpr int RPT = ST_ROWS * 64 / ST_THREADS;
constexpr int ST_G = ST_THREADS / 64;
constexpr int ST_MAX_BLOCKS = 4096;
int stage_grid(int rows, int cols) {
  const int t = (rows / ST_ROWS) * (cols / 64);
  return t < ST_MAX_BLOCKS ? t : ST_MAX_BLOCKS;
}

__device__ __forceinline__ float sum_splits(const float* ws, int ks, int64_t pstride, int64_t o) {
  float v = ws[o];
  for (int s = 1; s < ks; ++s) v += ws[s * pstride + o];
  return v;
}

template <class PP>
__global__ void __launch_bounds__(ST_THREADS)
    k_fwd_stage(EpiFwdStage e, const float* __restrict__ ws, int ks, int64_t pstride, int ldw,
                int tc, int ntile, const Ctl* __restrict__ ctl, int gate) {
  if (!gate_ok(ctl, gate)) return;
  using T = typename PP::T;
  const int tid = threadIdx.x, cl = tid & 63, rq = tid >> 6;
  const int64_t n = (e.mode == SM_STEP) ? ctl->n : 0;
  float* ent = e.tape + tape_slot(ctl, n) * e.ent;
  float* nxt = e.tape + tape_slot(ctl, n + 1) * e.ent;  // entry n + 1
  const float dt = (float)ctl->dt;
  const float atol = (float)ctl->abstol, rtol = (float)ctl->reltol;
  const int s = e.stage;
  T* X0 = (T*)e.X0;
  double sum = 0.0, flag = 0.0;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int r0 = (tile / tc) * ST_ROWS + RPT * rq, c = (tile % tc) * 64 + cl;
    if (c >= e.D) continue;
    const float bias = e.bias[c];
    float kv[RPT], uv[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = r0 + j;
      const int64_t idx = (int64_t)r * e.D + c;
      const float k = act_f(e.act, sum_splits(ws, ks, pstride, (int64_t)r * ldw + c) + bias);
      kv[j] = k;
      uv[j] = 0.f;
      const bool live = r < e.B;
      if (e.mode == SM_STEP) {
        if (s < 6) {
          float sm = cA[s + 1][s] * k;
#pragma unroll
          for (int i = 0; i < 5; ++i)
            if (i < s) sm += cA[s + 1][i] * ent[(1 + i) * e.sv + idx];
          uv[j] = ent[idx] + dt * sm;
        } else {
          float ut = cBT[6] * k;
#pragma unroll
          for (int i = 0; i < 6; ++i) ut += cBT[i] * ent[(1 + i) * e.sv + idx];
          ut *= dt;
          const float up = ent[idx], un = nxt[idx];
          const float sk = atol + fmaxf(fabsf(up), fabsf(un)) * rtol;
          if (live) {
            const double a = (double)(ut / sk);
            sum += a * a;
            if (isnan(un)) flag = 1.0;
          }
        }
      } else if (e.mode == SM_INIT_F0) {
        const float sk = atol + fabsf(ent[idx]) * rtol;
        if (live) { const double a = (double)(k / sk); sum += a * a; }
      } else if (e.mode == SM_INIT_F1) {
        const float f0 = ent[e.sv + idx];
        const float sk = atol + fabsf(ent[idx]) * rtol;
        if (live) {
          const double a = (double)((k - f0) / sk);
          sum += a * a;
          if (k != f0) flag = 1.0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int r = r0 + j;
      const int64_t idx = (int64_t)r * e.D + c;
      const float k = kv[j];
      if (e.mode == SM_STEP) {
        ent[(1 + s) * e.sv + idx] = k;
        if (s < 6) {
          put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, map_f(e.map, uv[j]));
          if (s == 5) nxt[idx] = uv[j];           // candidate tape[n+1].uprev
        } else {
          nxt[e.sv + idx] = k;                     // FSAL: tape[n+1].k1
        }
      } else if (e.mode == SM_INIT_F0) {
        ent[e.sv + idx] = k;
      } else {
        e.scratch[idx] = k;
      }
    }
  }
  if (e.part) block_reduce_store(sum, flag, e.part, blockIdx.x);
}

template <class PP>
__global__ void __launch_bounds__(ST_THREADS)
    k_bwd_stage(EpiBwdLast e, const float* __restrict__ ws, int ks, int64_t pstride, int ldw,
                int tc, int ntile, const Ctl* __restrict__ ctl, int gate) {
  if (!gate_ok(ctl, gate)) return;
  using T = typename PP::T;
  __shared__ float lt[2][64][ST_ROWS + 1];
  const int tid = threadIdx.x, cl = tid & 63, rq = tid >> 6;
  const float dt = (float)ctl->dt;
  const float atol = (float)ctl->abstol, rtol = (float)ctl->reltol;
  const int s = e.stage;
  const bool stage_out = e.mode == SM_STEP && s < 6;
  T* L = (T*)e.L;
  T* LT = (T*)e.LT;
  T* X0 = (T*)e.X0;
  T* X0T = (T*)e.X0T;
  // interpolation data of the next stage's time
  int tj = 0, ex = 1;
  float tdt = 0.f, w[7] = {0, 0, 0, 0, 0, 0, 0};
  if (stage_out) {
    tj = ctl->st_step[s + 1];
    ex = ctl->st_exact[s + 1];
    tdt = ctl->st_dt[s + 1];
#pragma unroll
    for (int i = 0; i < 7; ++i) w[i] = ctl->st_w[s + 1][i];
  }
  const float* te = e.tape + (int64_t)tj * e.ent;
  const float tnext = stage_out ? (float)(ctl->t + dC[s + 1] * ctl->dt) : 0.f;
  double sum = 0.0, flag = 0.0;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int tr0 = (tile / tc) * ST_ROWS, tcol = (tile % tc) * 64;
    const int r0 = tr0 + RPT * rq, c = tcol + cl;
    const bool cv = c < e.D;
    float kv[RPT], lsv[RPT], uv[RPT], xv[RPT], fuv[RPT];
    const bool bs = e.ubs != nullptr;  // BacksolveAdjoint: u integrated with lambda
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      kv[j] = lsv[j] = uv[j] = xv[j] = fuv[j] = 0.f;
      const int r = r0 + j;
      const int64_t idx = (int64_t)r * e.D + c;
      if (!cv) continue;
      // (a CNF's log p column: k = 0, its lambda stays constant)
      const bool zc = !e.cnf || c < e.dm;
      float g = zc ? sum_splits(ws, ks, pstride, (int64_t)r * ldw + c) : 0.f;
      if (e.map == 1) { const float u = e.Ucur[idx]; g *= 3.f * u * u; }
      const float k = zc ? -g : 0.f;
      kv[j] = e.mode == SM_PLAIN ? g : k;
      const bool live = r < e.B;
      if (bs && e.mode != SM_PLAIN) {
        // f(u_s) of this stage: u's own stage value and error terms
        const float fu = e.fu[idx];
        fuv[j] = fu;
        const float ub = e.ubs[idx];
        if (e.mode == SM_STEP && s == 6) {
          float ut = cBT[6] * fu;
#pragma unroll
          for (int i = 0; i < 6; ++i) ut += cBT[i] * e.ku[i * e.sv + idx];
          ut *= dt;
          const float un = e.ubs_new[idx];
          const float sk = atol + fmaxf(fabsf(ub), fabsf(un)) * rtol;
          if (live) {
            const double a = (double)(ut / sk);
            sum += a * a;
            if (isnan(un)) flag = 1.0;
          }
        } else if (e.mode == SM_INIT_F0 || e.mode == SM_INIT_F1) {
          const float d = e.mode == SM_INIT_F0 ? fu : fu - e.ku[idx];
          const float sk = atol + fabsf(ub) * rtol;
          if (live) {
            const double a = (double)(d / sk);
            sum += a * a;
            if (e.mode == SM_INIT_F1 && d != 0.f) flag = 1.0;
          }
        }
      }
      if (e.mode == SM_STEP) {
        if (s < 6) {
          float sm = cA[s + 1][s] * k;
#pragma unroll
          for (int i = 0; i < 5; ++i)
            if (i < s) sm += cA[s + 1][i] * e.kl[i * e.sv + idx];
          lsv[j] = e.lam[idx] + dt * sm;
          float u;
          if (bs) {
            float su = cA[s + 1][s] * fuv[j];
#pragma unroll
            for (int i = 0; i < 5; ++i)
              if (i < s) su += cA[s + 1][i] * e.ku[i * e.sv + idx];
            u = e.ubs[idx] + dt * su;
          } else {
            u = te[idx];
            if (!ex) {
              float sm2 = 0.f;
#pragma unroll
              for (int i = 0; i < 7; ++i) sm2 += w[i] * te[(1 + i) * e.sv + idx];
              u += tdt * sm2;
            }
          }
          uv[j] = u;
          xv[j] = map_f(e.map, u);
          if (e.cnf && c >= e.dm) xv[j] = e.tcat && c == e.dm && r < e.B ? tnext : 0.f;
        } else {
          float ut = cBT[6] * k;
#pragma unroll
          for (int i = 0; i < 6; ++i) ut += cBT[i] * e.kl[i * e.sv + idx];
          ut *= dt;
          const float lp = e.lam[idx], ln = e.lam_new[idx];
          const float sk = atol + fmaxf(fabsf(lp), fabsf(ln)) * rtol;
          if (live) {
            const double a = (double)(ut / sk);
            sum += a * a;
            if (isnan(ln)) flag = 1.0;
          }
        }
      } else if (e.mode == SM_INIT_F0) {
        const float sk = atol + fabsf(e.lam[idx]) * rtol;
        if (live) { const double a = (double)(k / sk); sum += a * a; }
      } else if (e.mode == SM_INIT_F1) {
        const float f0 = e.kl[idx];
        const float sk = atol + fabsf(e.lam[idx]) * rtol;
        if (live) {
          const double a = (double)((k - f0) / sk);
          sum += a * a;
          if (k != f0) flag = 1.0;
        }
      }
    }
    if (cv) {
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int64_t idx = (int64_t)(r0 + j) * e.D + c;
        if (e.mode == SM_PLAIN) {
          e.plain_out[idx] = kv[j];
        } else if (e.mode == SM_INIT_F0) {
          e.kl[idx] = kv[j];
          if (bs) e.ku[idx] = fuv[j];
        } else if (e.mode == SM_INIT_F1) {
          e.kl[7 * e.sv + idx] = kv[j];
        } else {
          e.kl[s * e.sv + idx] = kv[j];
          if (bs) e.ku[s * e.sv + idx] = fuv[j];
          if (s < 6) {
            if (s == 5) {
              e.lam_new[idx] = lsv[j];
              if (bs) e.ubs_new[idx] = uv[j];
            }
            e.Ucur[idx] = uv[j];
          }
        }
      }
    }
    if (stage_out) {
      // next-stage inputs lambda_s and x = map(u(t_s)): row-major (padded
      // columns get zeros) and transposed through LDS
      const bool lz = !e.cnf || c < e.dm;   // lambda columns the model sees
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int r = r0 + j;
        put<PP>(L + (int64_t)r * e.ldl + c, e.pl, lz ? lsv[j] : 0.f);
        put<PP>(X0 + (int64_t)r * e.ldx + c, e.px, xv[j]);
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        lt[0][cl][RPT * rq + j] = lz ? lsv[j] : 0.f;
        lt[1][cl][RPT * rq + j] = xv[j];
      }
      __syncthreads();
      const int cc = tid / ST_G, rr = RPT * (tid % ST_G);
      const int64_t col = tcol + cc;
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int r = tr0 + rr + j;
        put<PP>(LT + col * e.ldlt + r, e.plt, lt[0][cc][rr + j]);
        put<PP>(X0T + col * e.ldxt + r, e.pxt, lt[1][cc][rr + j]);
      }
    }
  }
  if (e.part) block_reduce_store(sum, flag, e.part, blockIdx.x);
}

void launch_fwd_stage_kernel(int prec, const EpiFwdStage& e, const float* ws, int ks,
                             int64_t pstride, int ldw, int rows, int cols, const Ctl* ctl,
                             int gate, hipStream_t s) {
  const int tc = cols / 64, nt = (rows / ST_ROWS) * tc;
  DISPATCH_PREC(prec, hipLaunchKernelGGL(k_fwd_stage<PP>, dim3(stage_grid(rows, cols)), dim3(ST_THREADS),
                                         0, s, e, ws, ks, pstride, ldw, tc, nt, ctl, gate));
}
void launch_bwd_stage_kernel(int prec, const EpiBwdLast& e, const float* ws, int ks,
                             int64_t pstride, int ldw, int rows, int cols, const Ctl* ctl,
                             int gate, hipStream_t s) {
  const int tc = cols / 64, nt = (rows / ST_ROWS) * tc;
  DISPATCH_PREC(prec, hipLaunchKernelGGL(k_bwd_stage<PP>, dim3(stage_grid(rows, cols)), dim3(ST_THREADS),
                                         0, s, e, ws, ks, pstride, ldw, tc, nt, ctl, gate));
}

// ---------------------------------------------------------------- elementwise
constexpr int EW_THREADS = 256;
constexpr int EW_MAX_BLOCKS = 1024;
int state_blocks(const StateGeom& sg) {
  int64_t n = (int64_t)sg.Bp * sg.D;
  int64_t b = (n + EW_THREADS - 1) / EW_THREADS;
  return (int)(b < EW_MAX_BLOCKS ? b : EW_MAX_BLOCKS);
}

// model input of state entry (r, c): map(u) for the model's columns; for a
// time-concatenated CNF model, column dm carries the stage time instead of
// log p (only for real rows: padded rows stay zero)
template <class PP>
__device__ __forceinline__ void put_x0(const StateGeom& sg, typename PP::T* X0, int64_t px,
                                       int r, int c, float u, double t) {
  if (c < sg.dm) put<PP>(X0 + (int64_t)r * sg.x0w + c, px, map_f(sg.map, u));
  else if (sg.tcat && c == sg.dm && r < sg.B) put<PP>(X0 + (int64_t)r * sg.x0w + c, px, (float)t);
}

template <class PP>
__global__ void k_fwd_init(StateGeom sg, const float* __restrict__ u0, float* tape,
                           typename PP::T* X0, int64_t px, const Ctl* ctl, double* part) {
  const double atol = ctl->abstol, rtol = ctl->reltol;
  double sum = 0.0;
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D), c = (int)(i % sg.D);
    const float u = r < sg.B ? u0[i] : 0.f;
    tape[i] = u;
    put_x0<PP>(sg, X0, px, r, c, u, ctl->t);
    if (r < sg.B) { const double a = (double)u / (atol + fabs((double)u) * rtol); sum += a * a; }
  }
  block_reduce_store(sum, 0.0, part, blockIdx.x);
}
void launch_fwd_init(const StateGeom& sg, const float* u0, float* tape, void* X0, int64_t px,
                     const Ctl* ctl, double* part, int nblk, hipStream_t s) {
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_fwd_init<PP>, dim3(nblk), dim3(EW_THREADS), 0, s,
                                            sg, u0, tape, (typename PP::T*)X0, px, ctl, part));
}

template <class PP>
__global__ void k_fwd_initdt_prep(StateGeom sg, const float* __restrict__ tape,
                                  typename PP::T* X0, int64_t px, const Ctl* ctl) {
  const float h = (float)(ctl->tdir * ctl->dt0);
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D), c = (int)(i % sg.D);
    const float u = tape[i] + h * tape[sg.sv + i];
    put_x0<PP>(sg, X0, px, r, c, u, ctl->t + ctl->tdir * ctl->dt0);
  }
}
void launch_fwd_initdt_prep(const StateGeom& sg, const float* tape, void* X0, int64_t px,
                            const Ctl* ctl, hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_fwd_initdt_prep<PP>, dim3(nb), dim3(EW_THREADS), 0,
                                            s, sg, tape, (typename PP::T*)X0, px, ctl));
}

// forward step begin: pending saves of the last accepted step, then the
// stage-2 input of the next attempt
template <class PP>
__global__ void k_fwd_step_begin(StateGeom sg, const float* __restrict__ tape,
                                 typename PP::T* X0, int64_t px,
                                 const double* __restrict__ save_times, float* out,
                                 const Ctl* ctl) {
  const int64_t n = (int64_t)sg.Bp * sg.D;
  const int64_t lo = ctl->save_lo, hi = ctl->save_hi;
  const double t = ctl->t;
  const int64_t cur = ctl->n;
  const float* ecur = tape + tape_slot(ctl, cur) * sg.ent;
  const float* es = tape + tape_slot(ctl, ctl->save_step) * sg.ent;
  const double tprev = ctl->tprev, dts = ctl->dt_step;
  const int64_t nreal = (int64_t)sg.B * sg.D;
  for (int64_t si = lo; si < hi; ++si) {
    const double ts = save_times[si];
    float w[7];
    const bool exact = (ts == t);
    if (!exact) {
      const double th = (ts - tprev) / dts, th2 = th * th;
      w[0] = (float)(th * (dR1[0] + th * (dR1[1] + th * (dR1[2] + th * dR1[3]))));
      for (int i = 0; i < 6; ++i) w[i + 1] = (float)(th2 * (dRI[i][0] + th * (dRI[i][1] + th * dRI[i][2])));
    }
    float* o = out + si * nreal;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nreal;
         i += (int64_t)gridDim.x * blockDim.x) {
      float v;
      if (exact) v = ecur[i];
      else {
        float sm = 0.f;
#pragma unroll
        for (int j = 0; j < 7; ++j) sm += w[j] * es[(1 + j) * sg.sv + i];
        v = es[i] + (float)dts * sm;
      }
      o[i] = v;
    }
  }
  if (ctl->done) return;
  const float h = (float)ctl->dt * cA[1][0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D), c = (int)(i % sg.D);
    const float u = ecur[i] + h * ecur[sg.sv + i];
    put_x0<PP>(sg, X0, px, r, c, u, ctl->t + dC[1] * ctl->dt);
  }
}
void launch_fwd_step_begin(const StateGeom& sg, const float* tape, void* X0, int64_t px,
                           const double* save_times, float* out, const Ctl* ctl,
                           hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_fwd_step_begin<PP>, dim3(nb), dim3(EW_THREADS), 0,
                                            s, sg, tape, (typename PP::T*)X0, px, save_times,
                                            out, ctl));
}

// ---------------------------------------------------------------- controllers
__device__ double block_sum_parts(const double* part, int n, double* flag_out) {
  __shared__ double rs[256], rf[256];
  double s = 0.0, f = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) { s += part[2 * i]; f = fmax(f, part[2 * i + 1]); }
  rs[threadIdx.x] = s;
  rf[threadIdx.x] = f;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rs[threadIdx.x] += rs[threadIdx.x + w];
      rf[threadIdx.x] = fmax(rf[threadIdx.x], rf[threadIdx.x + w]);
    }
    __syncthreads();
  }
  const double out = rs[0], fo = rf[0];
  __syncthreads();
  if (flag_out) *flag_out = fo;
  return out;
}

__device__ __forceinline__ void interp_weights(double th, float* w) {
  const double th2 = th * th;
  w[0] = (float)(th * (dR1[0] + th * (dR1[1] + th * (dR1[2] + th * dR1[3]))));
  for (int i = 0; i < 6; ++i) w[i + 1] = (float)(th2 * (dRI[i][0] + th * (dRI[i][1] + th * dRI[i][2])));
}

// SciMLBase sol(tau) on the forward record: exact stored state at a step
// time, else the interpolant of the bracketing step with dt = t_i - t_{i-1}
__device__ void set_interp(Ctl* c, int slot, double tau, const double* ft) {
  const int64_t nf = c->n_fwd;
  int64_t lo = 0, hi = nf + 1;  // first index with ft[i] >= tau
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (ft[mid] < tau) lo = mid + 1; else hi = mid;
  }
  int64_t i = lo;
  if (i <= nf && ft[i] == tau) { c->st_step[slot] = (int)i; c->st_exact[slot] = 1; return; }
  if (i > 0 && ft[i - 1] == tau) { c->st_step[slot] = (int)(i - 1); c->st_exact[slot] = 1; return; }
  if (i < 1) i = 1;
  if (i > nf) i = nf;
  const double dtj = ft[i] - ft[i - 1];
  c->st_step[slot] = (int)(i - 1);
  c->st_exact[slot] = 0;
  c->st_dt[slot] = (float)dtj;
  interp_weights((tau - ft[i - 1]) / dtj, c->st_w[slot]);
}

// loopheader + check_error for the next attempt; dt already holds the
// proposal (after accept) or the shrunk step (after reject).  C: the shared
// control block (Ctl) or one trajectory's registers (SlotCtl).
template <class C>
__device__ __forceinline__ void loopheader(C* c, double tstop) {
  c->iter += 1;
  double dt = c->dt;
  if (fabs(dt) > c->dtmax_abs) dt = c->tdir * c->dtmax_abs;
  const double dtmin_t = fmax(dspacing(c->t), c->dtmin);
  if (fabs(dt) < dtmin_t) dt = c->tdir * dtmin_t;
  if (c->adaptive) dt = c->tdir * fmin(fabs(dt), fabs(tstop - c->t));
  else dt = c->tdir * fmin(c->dt_fixed, fabs(tstop - c->t));
  c->dt = dt;
  if (isnan(dt)) { c->retcode = 3; c->done = 1; }
  else if (c->iter > c->maxiters) { c->retcode = 4; c->done = 1; }
  else if (c->adaptive && fabs(dt) <= c->dtmin &&
           (!c->last_accept || c->tdir * (c->t + dt) < c->tdir * tstop)) {
    c->retcode = 5; c->done = 1;
  }
}

__device__ __forceinline__ double snap_t(double t, double ttmp, double tstop) {
  return fabs(ttmp - tstop) < 100.0 * dspacing(fmax(t, tstop)) ? tstop : ttmp;
}

// accept/reject decision common to forward and adjoint; returns accepted
template <class C>
__device__ __forceinline__ bool pi_control(C* c, double eest) {
  c->eest = eest;
  if (!c->adaptive) return true;
  double q;
  if (eest == 0.0) q = 1.0 / QMAX;
  else {
    const double q11 = pow(eest, BETA1);
    c->q11 = q11;
    q = q11 / pow(c->qold, BETA2);
    q = fmax(1.0 / QMAX, fmin(1.0 / QMIN, q / GAMMA));
  }
  const bool acc = eest <= 1.0;
  if (acc) {
    if (q >= 1.0 && q <= 1.0) q = 1.0;
    c->qold = fmax(eest, QOLDINIT);
    double dtnew = c->dt / q;
    dtnew = c->tdir * fmin(c->dtmax_abs, fabs(dtnew));
    c->dtpropose = c->tdir * fmax(fabs(dtnew), c->dtmin);
  }
  return acc;
}

__global__ void k_initdt_a(Ctl* c, const double* p0, int n0, const double* px, int nx,
                           const double* pm, int nm, double nel) {
  double s0 = block_sum_parts(p0, n0, nullptr);
  double s1 = block_sum_parts(px, nx, nullptr);
  double sm = pm ? block_sum_parts(pm, nm, nullptr) : 0.0;
  if (threadIdx.x != 0) return;
  const double d0 = sqrt(s0 / nel), d1 = sqrt((s1 + sm) / nel);
  const double dtmin = nextafter(fmax(c->dtmin, dspacing(c->t)), INFINITY);
  const double smalldt = fmax(dtmin, 1e-6);
  double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
  dt0 = fmin(dt0, c->dtmax_abs);
  c->d0 = d0;
  c->d1 = d1;
  c->pad_ = 0;
  if (!(dt0 >= 10 * 2.220446049250313e-16)) { dt0 = smalldt; c->pad_ = 1; }
  c->dt0 = dt0;
}
void launch_initdt_a(Ctl* ctl, const double* p0, int n0, const double* px, int nx,
                     const double* pm, int nm, double n_elems, hipStream_t s) {
  hipLaunchKernelGGL(k_initdt_a, dim3(1), dim3(256), 0, s, ctl, p0, n0, px, nx, pm, nm, n_elems);
}

__global__ void k_initdt_b(Ctl* c, const double* px, int nx, const double* pm, int nm,
                           double nel, const double* save_times, int is_adj,
                           const double* ft) {
  double fl1 = 0.0, fl2 = 0.0;
  double s2 = block_sum_parts(px, nx, &fl1);
  double sm = pm ? block_sum_parts(pm, nm, &fl2) : 0.0;
  if (threadIdx.x != 0) return;
  if (c->has_dt_) {
    // dt given by the caller (adaptive initial dt or fixed step)
  } else if (c->pad_) {
    c->dt = c->tdir * c->dt0;
  } else {
    const double dt0 = c->dt0;
    const double dtmin = nextafter(fmax(c->dtmin, dspacing(c->t)), INFINITY);
    if (fmax(fl1, fl2) == 0.0) {
      c->dt = c->tdir * fmax(dtmin, 100 * dt0);
    } else {
      const double d2 = sqrt((s2 + sm) / nel) / dt0;
      const double m = fmax(c->d1, d2);
      double dt1;
      if (m <= 1e-15) dt1 = fmax(1e-6, dt0 * 1e-3);
      else dt1 = pow(10.0, -(2.0 + log10(m)) / 5.0);
      c->dt = c->tdir * fmax(dtmin, fmin(fmin(100 * dt0, dt1), c->dtmax_abs));
    }
  }
  c->nf += c->has_dt_ ? 0 : 1;
  c->dtpropose = c->dt;
  c->last_accept = 0;
  double tstop = c->t_end;
  if (!is_adj) {
    // saves at t0 (save_start)
    int64_t k = 0;
    while (k < c->n_save && c->tdir * save_times[k] <= c->tdir * c->t) ++k;
    c->save_lo = 0;
    c->save_hi = k;
    c->save_next = k;
    c->save_step = 0;
  } else {
    if (c->save_next >= 0) tstop = save_times[c->save_next];
  }
  loopheader(c, tstop);
  if (is_adj) {
    for (int s = 0; s < 7; ++s) set_interp(c, s, c->t + dC[s] * c->dt, ft);
  }
}
void launch_initdt_b(Ctl* ctl, const double* px, int nx, const double* pm, int nm,
                     double n_elems, const double* save_times, int is_adjoint,
                     const double* fwd_times, hipStream_t s) {
  hipLaunchKernelGGL(k_initdt_b, dim3(1), dim3(256), 0, s, ctl, px, nx, pm, nm, n_elems,
                     save_times, is_adjoint, fwd_times);
}

__global__ void k_fwd_controller(Ctl* c, const double* part, int np, const double* save_times,
                                 double* ft, double nel) {
  if (c->done) return;
  double flag = 0.0;
  const double s = block_sum_parts(part, np, &flag);
  if (threadIdx.x != 0) return;
  c->nf += 6;
  const double eest = c->adaptive ? sqrt(s / nel) : 0.0;
  if (c->adaptive && isnan(eest)) { c->retcode = 3; c->done = 1; c->save_lo = c->save_hi = c->save_next; return; }
  const bool acc = pi_control(c, eest);
  c->accepted = acc;
  if (acc) {
    c->naccept += 1;
    const double ttmp = c->t + c->dt;
    c->tprev = c->t;
    c->t = snap_t(c->t, ttmp, c->t_end);
    c->dt_step = c->dt;
    if (!c->adaptive) c->dtpropose = c->dt;
    c->save_step = c->n;
    c->n += 1;
    ft[tape_slot(c, c->n)] = c->t;
    int64_t k = c->save_next;
    c->save_lo = k;
    while (k < c->n_save && c->tdir * save_times[k] <= c->tdir * c->t) ++k;
    c->save_hi = k;
    c->save_next = k;
    c->last_accept = 1;
    if (flag != 0.0) { c->retcode = 6; c->done = 1; return; }
    if (c->t == c->t_end) { c->retcode = 1; c->done = 1; return; }
    c->dt = c->dtpropose;
  } else {
    c->nreject += 1;
    c->last_accept = 0;
    c->save_lo = c->save_hi = c->save_next;
    c->dt = c->dt / fmin(1.0 / QMIN, c->q11 / GAMMA);
  }
  loopheader(c, c->t_end);
}
void launch_fwd_controller(Ctl* ctl, const double* part, int n_part, const double* save_times,
                           double* fwd_times, double n_elems, hipStream_t s) {
  hipLaunchKernelGGL(k_fwd_controller, dim3(1), dim3(256), 0, s, ctl, part, n_part, save_times,
                     fwd_times, n_elems);
}
// ---------------------------------------------------------------- adjoint
template <class PP>
__device__ __forceinline__ void adj_prep_stage_u(const StateGeom& sg, const AdjBufs& ab,
                                                 int64_t i, float lam_s, float u, float t);
// time of interpolation slot `slot` of the next attempt (7: initial-dt probe)
__device__ __forceinline__ float slot_time(const Ctl* c, int slot) {
  return (float)(slot == 7 ? c->t + c->tdir * c->dt0 : c->t + dC[slot] * c->dt);
}
template <class PP>
__device__ __forceinline__ void adj_prep_stage(const StateGeom& sg, const AdjBufs& ab,
                                               const Ctl* c, int slot, int64_t i, float lam_s) {
  const float* te = ab.tape + (int64_t)c->st_step[slot] * sg.ent;
  float u = te[i];
  if (!c->st_exact[slot]) {
    float sm = 0.f;
#pragma unroll
    for (int j = 0; j < 7; ++j) sm += c->st_w[slot][j] * te[(1 + j) * sg.sv + i];
    u += c->st_dt[slot] * sm;
  }
  adj_prep_stage_u<PP>(sg, ab, i, lam_s, u, sg.tcat ? slot_time(c, slot) : 0.f);
}
// stage inputs (lambda_s row-major + transposed, x = map(u_s) likewise, u_s);
// a CNF's log p column (col == dm) feeds no lambda and, time-concatenated,
// the stage time t to the model
template <class PP>
__device__ __forceinline__ void adj_prep_stage_u(const StateGeom& sg, const AdjBufs& ab,
                                                 int64_t i, float lam_s, float u, float t) {
  using T = typename PP::T;
  const int r = (int)(i / sg.D), col = (int)(i % sg.D);
  const bool zc = col < sg.dm;
  const float lv = zc ? lam_s : 0.f;
  put<PP>((T*)ab.L + (int64_t)r * ab.ldl + col, ab.pl, lv);
  put<PP>((T*)ab.LT + (int64_t)col * ab.ldlt + r, ab.plt, lv);
  const float x = zc ? map_f(sg.map, u) : (sg.tcat && r < sg.B ? t : 0.f);
  put<PP>((T*)ab.X0 + (int64_t)r * ab.ldx + col, ab.px, x);
  put<PP>((T*)ab.X0T + (int64_t)col * ab.ldxt + r, ab.pxt, x);
  ab.U[i] = u;
}

template <class PP>
__global__ void k_adj_init(StateGeom sg, AdjBufs ab, const Ctl* c, double* part) {
  const double atol = c->abstol, rtol = c->reltol;
  const int jidx = c->jump_idx;
  double sum = 0.0;
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D);
    float l = 0.f;
    if (jidx >= 0 && r < sg.B) l = ab.dLdout[(int64_t)jidx * sg.B * sg.D + i];
    ab.lam[i] = l;
    adj_prep_stage<PP>(sg, ab, c, 0, i, l);
    if (r < sg.B) { const double a = (double)l / (atol + fabs((double)l) * rtol); sum += a * a; }
    if (ab.ubs) {  // backsolve: u starts from the forward's u(T)
      const float u = ab.U[i];
      ab.ubs[i] = u;
      if (r < sg.B) { const double a = (double)u / (atol + fabs((double)u) * rtol); sum += a * a; }
    }
  }
  block_reduce_store(sum, 0.0, part, blockIdx.x);
}
void launch_adj_init(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl, double* part,
                     hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_adj_init<PP>, dim3(nb), dim3(EW_THREADS), 0, s, sg,
                                            ab, ctl, part));
}

template <class PP>
__global__ void k_adj_initdt_prep(StateGeom sg, AdjBufs ab, const Ctl* c) {
  const float h = (float)(c->tdir * c->dt0);
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (ab.ubs) adj_prep_stage_u<PP>(sg, ab, i, ab.lam[i] + h * ab.kl[i], ab.ubs[i] + h * ab.ku[i], 0.f);
    else adj_prep_stage<PP>(sg, ab, c, 7, i, ab.lam[i] + h * ab.kl[i]);
  }
}
void launch_adj_initdt_prep(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl,
                            hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_adj_initdt_prep<PP>, dim3(nb), dim3(EW_THREADS), 0,
                                            s, sg, ab, ctl));
}

// adjoint step begin (idempotent, so speculative launches after the end are
// harmless): lambda = lambda_new + jump on accept, FSAL copy, mu commit into
// the other mu buffer, and the first stage input of the next attempt
template <class PP>
__global__ void k_adj_step_begin(StateGeom sg, AdjBufs ab, const Ctl* c, int nbs) {
  if ((int)blockIdx.x >= nbs) {
    if (!c->commit_mu) return;
    const int ms = c->gsel_mu_;
    const double* src = ms ? ab.mu : ab.mu + ab.P;
    double* dst = ms ? ab.mu + ab.P : ab.mu;
    const double h = c->dt_step;
    for (int64_t i = (blockIdx.x - nbs) * (int64_t)blockDim.x + threadIdx.x; i < ab.P;
         i += (int64_t)(gridDim.x - nbs) * blockDim.x)
      dst[i] = src[i] + h * (double)ab.accb[i];
    return;
  }
  const int64_t n = (int64_t)sg.Bp * sg.D;
  const int acc = c->last_accept, jidx = c->jump_idx, done = c->done, re = c->fsal_reeval;
  const float h = (float)c->dt * cA[1][0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)nbs * blockDim.x) {
    const int r = (int)(i / sg.D);
    float l = acc ? ab.lam_new[i] : ab.lam[i];
    if (acc && jidx >= 0 && r < sg.B) l += ab.dLdout[(int64_t)jidx * sg.B * sg.D + i];
    if (acc) { ab.lam[i] = l; ab.kl[i] = ab.kl[6 * sg.sv + i]; }
    if (ab.ubs) {  // backsolve: commit u; checkpoint = the forward's saved state
      float ub = acc ? ab.ubs_new[i] : ab.ubs[i];
      if (acc && jidx >= 0 && ab.ckpt && r < sg.B) ub = ab.fsaved[(int64_t)jidx * sg.B * sg.D + i];
      if (acc) { ab.ubs[i] = ub; ab.ku[i] = ab.ku[6 * sg.sv + i]; }
      if (done) continue;
      if (re) adj_prep_stage_u<PP>(sg, ab, i, l, ub, 0.f);
      else adj_prep_stage_u<PP>(sg, ab, i, l + h * ab.kl[i], ub + h * ab.ku[i], 0.f);
      continue;
    }
    if (done) continue;
    if (re) adj_prep_stage<PP>(sg, ab, c, 0, i, l);
    else adj_prep_stage<PP>(sg, ab, c, 1, i, l + h * (acc ? ab.kl[6 * sg.sv + i] : ab.kl[i]));
  }
}
int mu_blocks(int64_t P) {
  int64_t b = (P + EW_THREADS * 4 - 1) / (EW_THREADS * 4);
  return (int)(b < EW_MAX_BLOCKS ? (b < 1 ? 1 : b) : EW_MAX_BLOCKS);
}
void launch_adj_step_begin(const StateGeom& sg, const AdjBufs& ab, const Ctl* ctl, int gate,
                           hipStream_t s) {
  (void)gate;
  int nbs = state_blocks(sg), nbm = mu_blocks(ab.P);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_adj_step_begin<PP>, dim3(nbs + nbm),
                                            dim3(EW_THREADS), 0, s, sg, ab, ctl, nbs));
}

// mu part of the error norm: modes 0 step error, 1 initdt d1, 2 initdt d2
__global__ void k_mu_norm(int mode, AdjBufs ab, const Ctl* c, double* part, int gate) {
  if (!gate_ok(c, gate)) return;
  const double atol = c->abstol, rtol = c->reltol, dt = c->dt;
  const int ms = c->gsel_mu_;
  const double* mu = ms ? ab.mu + ab.P : ab.mu;
  const float* gcur = c->gsel ? ab.gf1 : ab.gf0;
  double sum = 0.0, flag = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ab.P;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double m = mu[i];
    double a;
    if (mode == 0) {
      const double mn = m + dt * (double)ab.accb[i];
      a = dt * (double)ab.acce[i] / (atol + fmax(fabs(m), fabs(mn)) * rtol);
    } else if (mode == 1) {
      a = (double)gcur[i] / (atol + fabs(m) * rtol);
    } else {
      a = ((double)ab.accb[i] - (double)gcur[i]) / (atol + fabs(m) * rtol);
      if (ab.accb[i] != gcur[i]) flag = 1.0;
    }
    sum += a * a;
  }
  block_reduce_store(sum, flag, part, blockIdx.x);
}
void launch_mu_norm(int mode, const AdjBufs& ab, const Ctl* ctl, double* part, int nblk,
                    int gate, hipStream_t s) {
  hipLaunchKernelGGL(k_mu_norm, dim3(nblk), dim3(EW_THREADS), 0, s, mode, ab, ctl, part, gate);
}

__global__ void k_adj_controller(Ctl* c, const double* pl, int nl, const double* pm, int nm,
                                 double nel, const double* cbt, const double* ft) {
  if (c->done) return;
  double flag = 0.0;
  const double sl = block_sum_parts(pl, nl, &flag);
  const double sm = block_sum_parts(pm, nm, nullptr);
  if (threadIdx.x != 0) return;
  c->nf += c->fsal_reeval ? 7 : 6;
  const double eest = c->adaptive ? sqrt((sl + sm) / nel) : 0.0;
  if (c->adaptive && isnan(eest)) { c->retcode = 3; c->done = 1; c->commit_mu = 0; c->last_accept = 0; c->jump_idx = -1; return; }
  const bool acc = pi_control(c, eest);
  c->accepted = acc;
  c->fsal_reeval = 0;
  c->jump_idx = -1;
  double tstop = c->save_next >= 0 ? cbt[c->save_next] : c->t_end;
  if (acc) {
    c->naccept += 1;
    const double ttmp = c->t + c->dt;
    c->tprev = c->t;
    c->t = snap_t(c->t, ttmp, tstop);
    c->dt_step = c->dt;
    if (!c->adaptive) c->dtpropose = c->dt;
    c->n += 1;
    c->last_accept = 1;
    c->commit_mu = 1;
    c->gsel_mu_ ^= 1;
    c->gsel ^= 1;
    if (c->save_next >= 0 && c->t == cbt[c->save_next]) {
      c->jump_idx = (int)c->save_next;
      c->save_next -= 1;
      c->fsal_reeval = 1;
    }
    if (flag != 0.0) { c->retcode = 6; c->done = 1; return; }
    if (c->t == c->t_end) { c->retcode = 1; c->done = 1; return; }
    c->dt = c->dtpropose;
  } else {
    c->nreject += 1;
    c->last_accept = 0;
    c->commit_mu = 0;
    c->dt = c->dt / fmin(1.0 / QMIN, c->q11 / GAMMA);
  }
  tstop = c->save_next >= 0 ? cbt[c->save_next] : c->t_end;
  loopheader(c, tstop);
  if (acc && c->has_pause && c->t == c->pause_t && !c->done) {
    // checkpoint reached: the host re-solves the forward interval below it
    // and resumes (k_adj_resume sets the stage interpolation data)
    c->paused = 1;
    c->done = 1;
    return;
  }
  for (int s = 0; s < 7; ++s) set_interp(c, s, c->t + dC[s] * c->dt, ft);
}

// Resume a paused adjoint on the forward interval just re-solved (n_fwd
// steps on the tape): the control block saved at the pause, the next
// checkpoint to pause at, and the forward-interpolation data of the next
// attempt's stages.  Stage 0 (the FSAL re-evaluation at the checkpoint time,
// where the cotangent jump has just been applied) reads the checkpoint state
// itself, which the host has placed in tape entry n_fwd + 1: at the
// checkpoint the adjoint is still on the interval above it, whose solution
// starts from that state.
__global__ void k_adj_resume(Ctl* c, const Ctl* saved, int64_t n_fwd, double pause_t,
                             int has_pause, const double* ft) {
  if (threadIdx.x != 0) return;
  *c = *saved;
  c->done = 0;
  c->paused = 0;
  c->n_fwd = n_fwd;
  c->pause_t = pause_t;
  c->has_pause = has_pause;
  for (int s = 0; s < 7; ++s) set_interp(c, s, c->t + dC[s] * c->dt, ft);
  c->st_step[0] = (int)(n_fwd + 1);
  c->st_exact[0] = 1;
}
void launch_adj_resume(Ctl* ctl, const Ctl* saved, int64_t n_fwd, double pause_t, int has_pause,
                       const double* fwd_times, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_resume, dim3(1), dim3(64), 0, s, ctl, saved, n_fwd, pause_t, has_pause,
                     fwd_times);
}
void launch_adj_controller(Ctl* ctl, const double* part, int n_lam, const double* part_mu,
                           int n_mu, double n_elems, const double* cb_times,
                           const double* fwd_times, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_controller, dim3(1), dim3(256), 0, s, ctl, part, n_lam, part_mu, n_mu,
                     n_elems, cb_times, fwd_times);
}

__global__ void k_set_stage_interp(Ctl* c, int slot, double tau, const double* ft) {
  if (threadIdx.x == 0) {
    double t = tau;
    if (slot == 7) t = c->t + c->tdir * c->dt0;
    set_interp(c, slot, t, ft);
  }
}
void launch_set_stage_interp(Ctl* ctl, int slot, double tau, const double* fwd_times,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_set_stage_interp, dim3(1), dim3(64), 0, s, ctl, slot, tau, fwd_times);
}

// fold: bias gradients g_b[o] = -sum_b delta^T[o][b] of every layer (one wave
// per bias row), and the fixed-order sum of the split-K partials of the
// narrow layers' weight gradients (one thread per entry); both folded into
// the mu accumulators like the weight-gradient epilogue does.
// fold_one on four consecutive entries (16-byte aligned)
__device__ __forceinline__ void fold_four(const EpiWgrad& e, float* gcur, float* gnext,
                                          int64_t idx, float4 g) {
  if (e.flags & WG_STORE_CUR) *(float4*)(gcur + idx) = g;
  if (e.flags & WG_STORE_PTR) *(float4*)(e.gptr + idx) = g;
  if (e.flags & WG_INIT) {
    const float4 g1 = *(const float4*)(gcur + idx);
    *(float4*)(e.accb + idx) = make_float4(e.cb1 * g1.x + e.cb * g.x, e.cb1 * g1.y + e.cb * g.y,
                                           e.cb1 * g1.z + e.cb * g.z, e.cb1 * g1.w + e.cb * g.w);
    *(float4*)(e.acce + idx) = make_float4(e.ce1 * g1.x + e.ce * g.x, e.ce1 * g1.y + e.ce * g.y,
                                           e.ce1 * g1.z + e.ce * g.z, e.ce1 * g1.w + e.ce * g.w);
  } else {
    if (e.flags & WG_B) {
      float4 a = *(const float4*)(e.accb + idx);
      a.x += e.cb * g.x; a.y += e.cb * g.y; a.z += e.cb * g.z; a.w += e.cb * g.w;
      *(float4*)(e.accb + idx) = a;
    }
    if (e.flags & WG_E) {
      float4 a = *(const float4*)(e.acce + idx);
      a.x += e.ce * g.x; a.y += e.ce * g.y; a.z += e.ce * g.z; a.w += e.ce * g.w;
      *(float4*)(e.acce + idx) = a;
    }
  }
  if (e.flags & WG_STORE_NEXT) *(float4*)(gnext + idx) = g;
}

template <class PP>
__global__ void k_fold(FoldArgs fa, EpiWgrad e, const Ctl* c, int gate, int nb_bias) {
  using T = typename PP::T;
  if (!gate_ok(c, gate)) return;
  float* gcur = c->gsel ? e.gf1 : e.gf0;
  float* gnext = c->gsel ? e.gf0 : e.gf1;
  if ((int)blockIdx.x < fa.nbw) {
    // bias rows of layers with per-row-tile column sums: one wave per row.
    // The row (and with it the layer) is wave-uniform: readfirstlane keeps
    // the per-layer argument lookups on scalar loads instead of per-lane loads
    // of the kernel-argument segment, one dependent round trip each
    const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int lane = threadIdx.x & 63;
    int l = 0, row = wave;
    for (; l < fa.nl; ++l) {
      if (!fa.bpart[l]) continue;
      if (row < fa.outs[l]) break;
      row -= fa.outs[l];
    }
    if (l >= fa.nl) return;
    const float* bp = fa.bpart[l] + row;
    float s = 0.f;
    for (int t = lane; t < fa.bpart_rt[l]; t += 64) s += bp[(int64_t)t * fa.bpart_ld[l]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) fold_one(e, gcur, gnext, fa.b0s[l] + row, -s);
    return;
  }
  if ((int)blockIdx.x < nb_bias) {
    // bias rows of the other layers (the transposed delta summed over the
    // batch): one block per row, eight independent partial sums per thread
    // (eight loads in flight), reduced in a fixed order (lanes, then waves)
    __shared__ float redw[4];
    int l = 0, row = (int)blockIdx.x - fa.nbw;
    for (; l < fa.nl; ++l) {
      if (fa.bpart[l]) continue;
      if (row < fa.outs[l]) break;
      row -= fa.outs[l];
    }
    if (l >= fa.nl) return;  // (block-uniform)
    const T* src = (const T*)fa.dT[l] + (int64_t)row * fa.ldt;
    float q[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int nt = (int)blockDim.x;
    int b = threadIdx.x;
    for (; b + 7 * nt < fa.Bp; b += 8 * nt) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = get<PP>(src + b + j * nt, fa.dplane[l]);
#pragma unroll
      for (int j = 0; j < 8; ++j) q[j] += v[j];
    }
    for (int j = 0; b < fa.Bp; b += nt, ++j) q[j] += get<PP>(src + b, fa.dplane[l]);
    float s = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) redw[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
      fold_one(e, gcur, gnext, fa.b0s[l] + row, -((redw[0] + redw[1]) + (redw[2] + redw[3])));
    return;
  }
  // split-K partial sums (at most 16 splits): a block takes 64 quads of four
  // consecutive entries of one layer; its four thread groups sum the splits
  // g, g + 4, g + 8, g + 12 of each quad (one round of independent loads) and
  // the groups are added in group order through LDS (deterministic)
  __shared__ float4 red[4][64];
  int bq = (int)blockIdx.x - nb_bias;
  const int ql = threadIdx.x & 63, grp = threadIdx.x >> 6;
  for (int k = 0; k < fa.nsk; ++k) {
    const int64_t n = fa.sk_n[k];
    const int nbk = (int)(((n + 3) / 4 + 63) / 64);
    if (bq < nbk) {
      const float* part = fa.sk_part[k];
      const int64_t w0 = fa.sk_w0[k];
      const int ks = fa.sk_split[k];
      const bool vec = ((w0 | n) & 3) == 0;   // 16-byte accesses
      const int64_t i = 4 * ((int64_t)bq * 64 + ql);
      float pv[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = grp + 4 * u;
#pragma unroll
        for (int c = 0; c < 4; ++c) pv[u][c] = 0.f;
        if (j < ks && i < n) {
          const float* src = part + (int64_t)j * n + i;
          if (vec) {
            const float4 v = *(const float4*)src;
            pv[u][0] = v.x; pv[u][1] = v.y; pv[u][2] = v.z; pv[u][3] = v.w;
          } else {
#pragma unroll
            for (int c = 0; c < 4; ++c)
              if (i + c < n) pv[u][c] = src[c];
          }
        }
      }
      float4 sv;
      sv.x = ((pv[0][0] + pv[1][0]) + pv[2][0]) + pv[3][0];
      sv.y = ((pv[0][1] + pv[1][1]) + pv[2][1]) + pv[3][1];
      sv.z = ((pv[0][2] + pv[1][2]) + pv[2][2]) + pv[3][2];
      sv.w = ((pv[0][3] + pv[1][3]) + pv[2][3]) + pv[3][3];
      red[grp][ql] = sv;
      __syncthreads();
      if (grp == 0 && i < n) {
        const float4 a = red[0][ql], b = red[1][ql], c2 = red[2][ql], d = red[3][ql];
        float4 t;
        t.x = (a.x + b.x) + (c2.x + d.x);
        t.y = (a.y + b.y) + (c2.y + d.y);
        t.z = (a.z + b.z) + (c2.z + d.z);
        t.w = (a.w + b.w) + (c2.w + d.w);
        if (vec) {
          fold_four(e, gcur, gnext, w0 + i, t);
        } else {
          const float tv[4] = {t.x, t.y, t.z, t.w};
          for (int c = 0; c < 4 && i + c < n; ++c) fold_one(e, gcur, gnext, w0 + i + c, tv[c]);
        }
      }
      return;
    }
    bq -= nbk;
  }
}
void launch_fold(int prec, const FoldArgs& fa_in, const EpiWgrad& tmpl, const Ctl* ctl, int gate,
                 hipStream_t s) {
  // blocks: wave-per-row bias rows of the layers with column-sum partials,
  // then one block per bias row of the other layers, then the split-K sums
  FoldArgs fa = fa_in;
  int rows_w = 0, rows_b = 0;
  for (int l = 0; l < fa.nl; ++l) (fa.bpart[l] ? rows_w : rows_b) += fa.outs[l];
  fa.nbw = (rows_w * 64 + 255) / 256;
  const int nb_bias = fa.nbw + rows_b;
  int64_t nsk = 0;  // blocks of 64 quads per split-K layer
  for (int k = 0; k < fa.nsk; ++k) nsk += ((fa.sk_n[k] + 3) / 4 + 63) / 64;
  const int nb = nb_bias + (int)nsk;
  DISPATCH_PREC(prec, hipLaunchKernelGGL(k_fold<PP>, dim3(nb), dim3(256), 0, s, fa, tmpl, ctl,
                                         gate, nb_bias));
}

// plain input preparation for node_rhs / node_vjp
template <class PP>
__global__ void k_prep_plain(StateGeom sg, const float* u, typename PP::T* X0,
                             typename PP::T* X0T, float* Ucur, int ldx, int ldxt, int64_t px,
                             int64_t pxt) {
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D), c = (int)(i % sg.D);
    const float v = r < sg.B ? u[i] : 0.f;
    const float x = map_f(sg.map, v);
    put<PP>(X0 + (int64_t)r * ldx + c, px, x);
    if (X0T) put<PP>(X0T + (int64_t)c * ldxt + r, pxt, x);
    if (Ucur) Ucur[i] = v;
  }
}
void launch_prep_plain(const StateGeom& sg, const float* u, void* X0, void* X0T, float* Ucur,
                       int ldx, int ldxt, int64_t px, int64_t pxt, hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_prep_plain<PP>, dim3(nb), dim3(EW_THREADS), 0, s, sg,
                                            u, (typename PP::T*)X0, (typename PP::T*)X0T, Ucur,
                                            ldx, ldxt, px, pxt));
}
// zero-fill of a 16-byte aligned region (memory a captured graph may clear:
// a kernel, not hipMemsetAsync)
__global__ void k_fill_zero(u32x4* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = u32x4{0u, 0u, 0u, 0u};
}
void launch_fill_zero(void* p, int64_t bytes, hipStream_t s) {
  const int64_t n = bytes / 16;
  if (n <= 0) return;
  const int64_t nb = (n + EW_THREADS - 1) / EW_THREADS;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)(nb < 2048 ? nb : 2048)), dim3(EW_THREADS), 0, s,
                     (u32x4*)p, n);
}
template <class PP>
__global__ void k_prep_lambda(StateGeom sg, const float* lam, typename PP::T* L,
                              typename PP::T* LT, int ldl, int ldlt, int64_t pl, int64_t plt) {
  const int64_t n = (int64_t)sg.Bp * sg.D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / sg.D), c = (int)(i % sg.D);
    const float v = r < sg.B ? lam[i] : 0.f;
    put<PP>(L + (int64_t)r * ldl + c, pl, v);
    if (LT) put<PP>(LT + (int64_t)c * ldlt + r, plt, v);
  }
}
void launch_prep_lambda(const StateGeom& sg, const float* lam, void* L, void* LT, int ldl,
                        int ldlt, int64_t pl, int64_t plt, hipStream_t s) {
  int nb = state_blocks(sg);
  DISPATCH_PREC(sg.prec, hipLaunchKernelGGL(k_prep_lambda<PP>, dim3(nb), dim3(EW_THREADS), 0, s,
                                            sg, lam, (typename PP::T*)L, (typename PP::T*)LT,
                                            ldl, ldlt, pl, plt));
}

// gather real rows of n state vectors (stride src_stride floats) into a dense
// D x B x n array
__global__ void k_gather(StateGeom sg, const float* src, int64_t stride, float* dst, int64_t n) {
  const int64_t per = (int64_t)sg.B * sg.D;
  const int64_t tot = per * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i / per, k = i % per;
    dst[i] = src[j * stride + k];
  }
}
void launch_gather(const StateGeom& sg, const float* src, int64_t src_stride_floats, float* dst,
                   int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3(EW_MAX_BLOCKS), dim3(EW_THREADS), 0, s, sg, src,
                     src_stride_floats, dst, n);
}

#include "ragged_kernels.h"

}  // namespace nodehip
