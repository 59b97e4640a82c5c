// Per-trajectory ("ragged") solves: many independent trajectories of one
// model, each its own ODE problem with its own tspan end, save times, error
// norm (RMS over its own D entries) and PI-controlled steps -- what N separate
// solves of batch 1 compute -- run by one persistent launch.
//
// Included by kernels.hip (shares the tableau, controller and activation
// helpers).
//
// Layout of the work: a wave owns RG_SLOTS = 16 trajectory slots, the 16
// columns of v_mfma_f32_16x16x4f32 tiles whose rows are features.  Lane l
// serves slot l % 16 and features 16*t + 4*(l/16) + v (v = 0..3) of every
// 16-feature tile t, which is both the accumulator layout of a tile and the
// operand layout the next layer needs (k-index l/16 of MFMA step v takes
// feature 4*(l/16) + v), so activations stay in registers through the whole
// Dense chain.  Weights sit in LDS once per workgroup as FP32 [out][in + 4]
// rows: the forward reads 16-byte rows (W[o][i..i+3]), the transposed
// products of the adjoint read W[o..o+3][i] column entries from the same copy.
// Each slot runs its own state machine (initial f0, initial-dt f1, stages
// 2..7 of an attempt); every round evaluates the right-hand side once for all
// 16 slots, whatever phase each is in.  A slot whose trajectory has finished
// takes the next one from a global counter, so long and short trajectories
// share the machine until the queue is empty (the adjoint starts each slot
// from a host-computed, cost-balanced order and may run fewer than 16 slots
// per wave: RaggedAdjArgs).  Every trajectory's arithmetic depends only on
// its own column, so results do not depend on scheduling.

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Global-memory pointers.  The kernels get their buffers inside argument
// structs, whose pointers the compiler leaves generic: their accesses become
// FLAT instructions, which count against the LDS / scalar counter too, so a
// wave can only wait for one of their results by draining every LDS and
// scalar access in flight.  The hot loops go through these instead.
template <class T> using gp = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ gp<T> glb(T* p) { return (gp<T>)p; }

struct SlotCtl {
  double t, dt, tprev, dtpropose, qold, q11, eest, tdir, dtmax_abs, dtmin, t_end, dt_fixed;
  int64_t iter, maxiters;
  int adaptive, retcode, done, last_accept;
};

__device__ __forceinline__ f32x4 rg_mfma(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// BF16 operands for v_mfma_f32_16x16x16_bf16: a lane's 4 values of a 16-wide
// k-tile (k = 4 (l / 16) + v) are that MFMA's k-slice for the lane, the same
// layout as the four FP32 steps above, so one instruction replaces four
typedef short rg_s4 __attribute__((ext_vector_type(4)));
typedef __bf16 rg_b4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ rg_s4 rg_bf4(const f32x4& v) {
  const rg_b4 b = {(__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
  return __builtin_bit_cast(rg_s4, b);
}
__device__ __forceinline__ f32x4 rg_mfma_bf(rg_s4 a, rg_s4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// z[ot] = W x for output tiles ot < nto (no bias); x: nti input tiles
template <int NTI, int NTO>
__device__ __forceinline__ void rg_fwd_gemm(const float* W, int pitch, const f32x4 (&x)[NTI],
                                            int nti, int nto, f32x4 (&z)[NTO], int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int ot = 0; ot < NTO; ++ot) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (ot < nto) {
      const float* wrow = W + (16 * ot + r) * pitch + 4 * g;
#pragma unroll
      for (int it = 0; it < NTI; ++it) {
        if (it < nti) {
          const f32x4 a = *(const f32x4*)(wrow + 16 * it);
          acc = rg_mfma(a.x, x[it].x, acc);
          acc = rg_mfma(a.y, x[it].y, acc);
          acc = rg_mfma(a.z, x[it].z, acc);
          acc = rg_mfma(a.w, x[it].w, acc);
        }
      }
    }
    z[ot] = acc;
  }
}

// out[it] = W^T d for input tiles it < nti; d: nto output-feature tiles
template <int NTI, int NTO>
__device__ __forceinline__ void rg_bwd_gemm(const float* W, int pitch, const f32x4 (&d)[NTO],
                                            int nti, int nto, f32x4 (&out)[NTI], int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int it = 0; it < NTI; ++it) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (it < nti) {
#pragma unroll
      for (int ot = 0; ot < NTO; ++ot) {
        if (ot < nto) {
          const float* wc = W + (16 * ot + 4 * g) * pitch + 16 * it + r;
          acc = rg_mfma(wc[0], d[ot].x, acc);
          acc = rg_mfma(wc[pitch], d[ot].y, acc);
          acc = rg_mfma(wc[2 * pitch], d[ot].z, acc);
          acc = rg_mfma(wc[3 * pitch], d[ot].w, acc);
        }
      }
    }
    out[it] = acc;
  }
}

// one Dense layer, tile by tile: y[ot] = act(W x + b) for ot < nto
template <int NTI, int NTO>
__device__ __forceinline__ void rg_layer(const float* W, int pitch, const float* bias, int act,
                                         const f32x4 (&x)[NTI], int nti, int nto,
                                         f32x4 (&y)[NTO], int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int ot = 0; ot < NTO; ++ot) {
    if (ot < nto) {
      f32x4 acc = *(const f32x4*)(bias + 16 * ot + 4 * g);
      const float* wrow = W + (16 * ot + r) * pitch + 4 * g;
#pragma unroll
      for (int it = 0; it < NTI; ++it) {
        if (it < nti) {
          const f32x4 a = *(const f32x4*)(wrow + 16 * it);
          acc = rg_mfma(a.x, x[it].x, acc);
          acc = rg_mfma(a.y, x[it].y, acc);
          acc = rg_mfma(a.z, x[it].z, acc);
          acc = rg_mfma(a.w, x[it].w, acc);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) y[ot][v] = act_f(act, acc[v]);
    }
  }
}

// The model's forward pass for the 16 slot columns: x (DT state tiles, input
// map applied) -> f (DT tiles).  One rolled loop body over the layers (the
// output is copied into place after each layer) keeps the code small enough
// for the instruction cache.
template <int DT>
__device__ __forceinline__ void rg_mlp(const RaggedModel& m, const float* lds,
                                       const f32x4 (&x)[DT], f32x4 (&f)[DT], int lane) {
  f32x4 hin[RG_HT], hout[RG_HT];
#pragma unroll
  for (int t = 0; t < RG_HT; ++t) hin[t] = t < DT ? x[t < DT ? t : 0] : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int l = 0; l < m.L; ++l) {
    rg_layer<RG_HT, RG_HT>(lds + m.lw[l], m.pitch[l], lds + m.lb[l], m.act[l], hin, m.ti[l],
                           m.to[l], hout, lane);
#pragma unroll
    for (int t = 0; t < RG_HT; ++t) hin[t] = hout[t];
  }
#pragma unroll
  for (int t = 0; t < DT; ++t) f[t] = hin[t];
}

// weights of every layer into LDS rows [out_p][in_p + 4], bias [out_p];
// padded rows / columns zero (they then contribute nothing)
__device__ void rg_load_weights(const RaggedModel& m, float* lds) {
  for (int i = threadIdx.x; i < m.lds_floats; i += blockDim.x) lds[i] = 0.f;
  __syncthreads();
  for (int l = 0; l < m.L; ++l) {
    const int in = m.in[l], out = m.out[l], pitch = m.pitch[l];
    const float* p = m.p + m.w0[l];
    for (int e = threadIdx.x; e < in * out; e += blockDim.x) {
      const int i = e / out, o = e - i * out;     // Julia W[o, i] at i*out + o
      lds[m.lw[l] + o * pitch + i] = p[e];
    }
    for (int o = threadIdx.x; o < out; o += blockDim.x) lds[m.lb[l] + o] = m.p[m.b0[l] + o];
  }
  __syncthreads();
}

// sum over the 4 lanes (l, l^16, l^32, l^48) that share a slot
__device__ __forceinline__ double rg_slot_sum(double v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}
__device__ __forceinline__ int rg_bcast(int v, int lane) { return __shfl(v, lane & 15); }

__device__ __forceinline__ float rg_map(int m, float u) { return m == 1 ? u * u * u : u; }

// sum_{j<s} A[s][j] k_j of tile t, newest stage first (the order the batched
// stage epilogues use, so u_new is formed identically for the k7 input and
// the accepted state)
template <int DT>
__device__ __forceinline__ f32x4 rg_stage_sum(const f32x4 (&k)[7][DT], int s, int t) {
  f32x4 sm = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if (j == s - 1) sm = cA[s][j] * k[j][t];
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if (j < s - 1) sm += cA[s][j] * k[j][t];
  return sm;
}

// ---------------------------------------------------------------- forward
// one tape record: start time, link to the trajectory's previous record, the
// state at that time and the nk stages of the step taken from it
template <int DT>
__device__ __forceinline__ int rg_record(const RaggedFwdArgs& a, double t, int prev,
                                         const f32x4 (&uu)[DT], const f32x4 (&k)[7][DT],
                                         int nk, int lane) {
  const int g = lane >> 4;
  int idx = 0;
  if (g == 0) idx = atomicAdd(a.rec_count, 1);
  idx = rg_bcast(idx, lane);
  if (idx < a.rec_cap) {
    if (g == 0) {
      a.rec_t[idx] = t;
      a.rec_prev[idx] = prev;
    }
    float* rd = a.rec_data + (int64_t)idx * 8 * a.Dp;
#pragma unroll
    for (int t2 = 0; t2 < DT; ++t2) {
      *(f32x4*)(rd + 16 * t2 + 4 * g) = uu[t2];
#pragma unroll
      for (int j = 0; j < 7; ++j)
        if (j < nk) *(f32x4*)(rd + (1 + j) * a.Dp + 16 * t2 + 4 * g) = k[j][t2];
    }
  }
  return idx;
}

template <int DT>
__global__ void __launch_bounds__(256) k_ragged_fwd(RaggedFwdArgs a) {
  extern __shared__ __attribute__((aligned(16))) float rg_lds[];
  const RaggedModel& m = a.m;
  rg_load_weights(m, rg_lds);
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int D = a.D;
  const float atol = (float)a.abstol, rtol = (float)a.reltol;

  // slot state (identical in the slot's 4 lanes)
  int traj = -1, phase = 0, stage = 0, pad = 0, last_rec = -1, nrec = 0;
  bool need_start = true;
  SlotCtl c{};
  f32x4 u[DT], k[7][DT];
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    u[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 7; ++j) k[j][t] = u[t];
  }
  int64_t sidx = 0, send = 0, nf = 0, nacc = 0, nrej = 0;
  double dt0 = 0.0, d1 = 0.0;

  for (;;) {
    if (need_start) {
      need_start = false;
      int next = 0;
      if (g == 0) next = atomicAdd(a.queue, 1);
      next = rg_bcast(next, lane);
      traj = next < a.n ? next : -1;
      if (traj >= 0) {
#pragma unroll
        for (int t = 0; t < DT; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int f = 16 * t + 4 * g + v;
            u[t][v] = f < D ? a.u0[(int64_t)traj * D + f] : 0.f;
          }
        c = SlotCtl{};
        c.t = a.t0;
        c.t_end = a.t1[traj];
        c.tdir = c.t_end > c.t ? 1.0 : -1.0;
        c.dtmax_abs = a.dtmax > 0 ? a.dtmax : fabs(c.t_end - c.t);
        c.dtmin = a.dtmin > 0 ? a.dtmin : fmax(dspacing(a.t0), dspacing(c.t_end));
        c.q11 = 1.0;
        c.maxiters = a.maxiters;
        c.adaptive = a.adaptive;
        c.dt_fixed = a.dt_init;
        c.qold = QOLDINIT;
        sidx = a.save_off[traj];
        send = a.save_off[traj + 1];
        while (sidx < send && c.tdir * a.save_t[sidx] <= c.tdir * c.t) {  // saves at t0
#pragma unroll
          for (int t = 0; t < DT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int f = 16 * t + 4 * g + v;
              if (f < D) a.out[sidx * D + f] = u[t][v];
            }
          ++sidx;
        }
        last_rec = -1;
        nrec = 0;
        nf = nacc = nrej = 0;
        phase = 0;
        stage = 0;
      }
    }
    if (!__any(traj >= 0)) break;
    // ---- right-hand-side input of this round
    f32x4 x[DT];
    const float h = (float)c.dt;
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      f32x4 v = u[t];
      if (traj < 0) v = f32x4{0.f, 0.f, 0.f, 0.f};
      else if (phase == 1) v += (float)(c.tdir * dt0) * k[0][t];
      else if (phase == 2) v += h * rg_stage_sum(k, stage, t);
#pragma unroll
      for (int q = 0; q < 4; ++q) x[t][q] = rg_map(m.map, v[q]);
    }
    f32x4 fo[DT];
    rg_mlp<DT>(m, rg_lds, x, fo, lane);
    if (traj < 0) continue;
    ++nf;
    bool fin = false;
    // ---- consume
    if (phase == 0) {  // f0 = f(u0); initial-dt d0, d1
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int t = 0; t < DT; ++t) {
        k[0][t] = fo[t];
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (16 * t + 4 * g + v < D) {
            const float sk = atol + fabsf(u[t][v]) * rtol;
            const double a0 = (double)u[t][v] / (a.abstol + fabs((double)u[t][v]) * a.reltol);
            const double a1 = (double)(fo[t][v] / sk);
            s0 += a0 * a0;
            s1 += a1 * a1;
          }
      }
      s0 = rg_slot_sum(s0);
      s1 = rg_slot_sum(s1);
      if (a.dt_init > 0.0) {  // caller's dt: no heuristic
        c.dt = c.tdir * a.dt_init;
        c.dtpropose = c.dt;
        c.last_accept = 0;
        loopheader(&c, c.t_end);
        phase = 2;
        stage = 1;
        fin = c.done;
      } else {
        const double d0 = sqrt(s0 / D);
        d1 = sqrt(s1 / D);
        const double dtmin = nextafter(fmax(c.dtmin, dspacing(c.t)), INFINITY);
        const double smalldt = fmax(dtmin, 1e-6);
        dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
        dt0 = fmin(dt0, c.dtmax_abs);
        pad = 0;
        if (!(dt0 >= 10 * 2.220446049250313e-16)) { dt0 = smalldt; pad = 1; }
        phase = 1;
      }
    } else if (phase == 1) {  // f1 at u0 + dt0 f0: initial dt
      double s2 = 0.0, fl = 0.0;
#pragma unroll
      for (int t = 0; t < DT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (16 * t + 4 * g + v < D) {
            const float sk = atol + fabsf(u[t][v]) * rtol;
            const double a2 = (double)((fo[t][v] - k[0][t][v]) / sk);
            s2 += a2 * a2;
            if (fo[t][v] != k[0][t][v]) fl = 1.0;
          }
      s2 = rg_slot_sum(s2);
      fl = rg_slot_sum(fl);
      const double dtmin = nextafter(fmax(c.dtmin, dspacing(c.t)), INFINITY);
      if (pad) {
        c.dt = c.tdir * dt0;
      } else if (fl == 0.0) {
        c.dt = c.tdir * fmax(dtmin, 100 * dt0);
      } else {
        const double d2 = sqrt(s2 / D) / dt0;
        const double mx = fmax(d1, d2);
        const double dt1 = mx <= 1e-15 ? fmax(1e-6, dt0 * 1e-3) : pow(10.0, -(2.0 + log10(mx)) / 5.0);
        c.dt = c.tdir * fmax(dtmin, fmin(fmin(100 * dt0, dt1), c.dtmax_abs));
      }
      c.dtpropose = c.dt;
      c.last_accept = 0;
      loopheader(&c, c.t_end);
      phase = 2;
      stage = 1;
      fin = c.done;
    } else {  // stage `stage` of an attempt: k[stage] = f(...)
#pragma unroll
      for (int j = 1; j < 7; ++j)
        if (j == stage)
#pragma unroll
          for (int t = 0; t < DT; ++t) k[j][t] = fo[t];
      if (stage < 6) {
        ++stage;
      } else {
        // k7 = f(u_new), u_new = u + dt sum_j b_j k_j (b = row 7 of A)
        f32x4 un[DT];
        double sum = 0.0, fl = 0.0;
#pragma unroll
        for (int t = 0; t < DT; ++t) {
          un[t] = u[t] + h * rg_stage_sum(k, 6, t);
          f32x4 ut = cBT[6] * k[6][t];
#pragma unroll
          for (int j = 0; j < 6; ++j) ut += cBT[j] * k[j][t];
          ut *= h;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (16 * t + 4 * g + v < D) {
              const float sk = atol + fmaxf(fabsf(u[t][v]), fabsf(un[t][v])) * rtol;
              const double e = (double)(ut[v] / sk);
              sum += e * e;
              if (isnan(un[t][v])) fl = 1.0;
            }
        }
        sum = rg_slot_sum(sum);
        fl = rg_slot_sum(fl);
        const double eest = c.adaptive ? sqrt(sum / D) : 0.0;
        if (c.adaptive && isnan(eest)) {
          c.retcode = 3;
          c.done = 1;
        } else if (pi_control(&c, eest)) {
          ++nacc;
          last_rec = rg_record<DT>(a, c.t, last_rec, u, k, 7, lane);
          ++nrec;
          const double tprev = c.t, dts = c.dt;
          c.t = snap_t(c.t, c.t + c.dt, c.t_end);
          while (sidx < send && c.tdir * a.save_t[sidx] <= c.tdir * c.t) {  // saves in (tprev, t]
            const double ts = a.save_t[sidx];
            float w[7];
            const bool exact = ts == c.t;
            if (!exact) interp_weights((ts - tprev) / dts, w);
#pragma unroll
            for (int t = 0; t < DT; ++t)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                float val = un[t][v];
                if (!exact) {
                  float sm = 0.f;
#pragma unroll
                  for (int j = 0; j < 7; ++j) sm += w[j] * k[j][t][v];
                  val = u[t][v] + (float)dts * sm;
                }
                const int f = 16 * t + 4 * g + v;
                if (f < D) a.out[sidx * D + f] = val;
              }
            ++sidx;
          }
#pragma unroll
          for (int t = 0; t < DT; ++t) { u[t] = un[t]; k[0][t] = k[6][t]; }
          if (!c.adaptive) c.dtpropose = c.dt;
          c.last_accept = 1;
          if (fl != 0.0) { c.retcode = 6; c.done = 1; }
          else if (c.t == c.t_end) { c.retcode = 1; c.done = 1; }
          else {
            c.dt = c.dtpropose;
            loopheader(&c, c.t_end);
          }
        } else {
          ++nrej;
          c.last_accept = 0;
          c.dt = c.dt / fmin(1.0 / QMIN, c.q11 / GAMMA);
          loopheader(&c, c.t_end);
        }
        stage = 1;
        fin = c.done;
      }
    }
    if (fin) {
      last_rec = rg_record<DT>(a, c.t, last_rec, u, k, 1, lane);  // final state, k1 = f(u_end)
      ++nrec;
      if (g == 0) {
        a.retcode[traj] = c.retcode;
        a.last_rec[traj] = last_rec;
        int* ts = a.traj_stats + 4 * (int64_t)traj;
        ts[0] = (int)nf; ts[1] = (int)nacc; ts[2] = (int)nrej; ts[3] = nrec;
        atomicAdd(a.stats + 0, (unsigned long long)nf);
        atomicAdd(a.stats + 1, (unsigned long long)nacc);
        atomicAdd(a.stats + 2, (unsigned long long)nrej);
      }
      need_start = true;
    }
  }
}

void launch_ragged_fwd(const RaggedFwdArgs& a, int grid, hipStream_t s) {
  const size_t lds = (size_t)a.m.lds_floats * sizeof(float);
  if (a.Dp <= 16) {
    (void)hipFuncSetAttribute((const void*)k_ragged_fwd<1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_ragged_fwd<1>, dim3(grid), dim3(256), lds, s, a);
  } else {
    (void)hipFuncSetAttribute((const void*)k_ragged_fwd<2>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_ragged_fwd<2>, dim3(grid), dim3(256), lds, s, a);
  }
}

// ---------------------------------------------------------------- adjoint
// y = act(z) and act'(z) of a 4-entry vector, the activation chosen once per
// vector (a per-entry switch puts a chain of branches around every entry)
__device__ __forceinline__ void rg_act4(int act, const f32x4& z, f32x4& y, f32x4& sd) {
#define RG_ACT_CASE(A)                                                   \
  case A:                                                                \
    _Pragma("unroll") for (int v = 0; v < 4; ++v) {                      \
      y[v] = act_f(A, z[v]);                                             \
      sd[v] = act_d(A, z[v], y[v]);                                      \
    }                                                                    \
    break;
  switch (act) {
    RG_ACT_CASE(1)
    RG_ACT_CASE(2)
    RG_ACT_CASE(3)
    RG_ACT_CASE(4)
    default:
      y = z;
      sd = f32x4{1.f, 1.f, 1.f, 1.f};
  }
#undef RG_ACT_CASE
}

// One VJP for the 16 slot columns: x0 = map(u) (first ti[0] tiles), lam (first
// to[L-1] tiles).  Stores the layer inputs h_l and the pre-activation
// gradients delta_l of this evaluation in the slot's stage block `blk`
// (act' is parked in the wave's LDS area `act_lds` during the forward sweep,
// or in the block's delta area when m.act_stride is 0) and returns
// g = W_0^T delta_0 (the input-map derivative is applied by the caller).
// BF: the products on BF16 operands (rounded to nearest), FP32 accumulation.
template <bool BF>
__device__ __forceinline__ void rg_vjp(const RaggedModel& m, const float* lds, float* act_lds,
                                       const f32x4 (&x0)[RG_HT], const f32x4 (&lam)[RG_HT],
                                       gp<float> blk, f32x4 (&gout)[RG_HT], int lane) {
  const int r = lane & 15, g4 = 4 * (lane >> 4);
  const int L = m.L;
  // one rolled loop body per sweep (the layer's input is copied into place
  // after each layer): the unrolled-over-layers form did not fit the
  // instruction cache
  f32x4 hin[RG_HT], hout[RG_HT];
#pragma unroll
  for (int t = 0; t < RG_HT; ++t) {
    hin[t] = x0[t];
    if (t < m.ti[0]) *(gp<f32x4>)(blk + m.sh[0] + 16 * t + g4) = x0[t];
  }
#pragma unroll 1
  for (int l = 0; l < L; ++l) {
    const float* W = lds + m.lw[l];
    const float* bias = lds + m.lb[l];
    const int pitch = m.pitch[l], nti = m.ti[l], nto = m.to[l], act = m.act[l];
    rg_s4 hb[RG_HT];
    if (BF) {
#pragma unroll
      for (int t = 0; t < RG_HT; ++t) hb[t] = rg_bf4(hin[t]);
    }
    // a hidden tanh layer of 8 x 8 tiles with act' parked in LDS (the usual
    // shape): one straight-line block, so the compiler can run a tile's
    // activation and stores under the next tile's reads and MFMAs
    if (nti == RG_HT && nto == RG_HT && act == 1 && m.act_stride && l < L - 1) {
#pragma unroll
      for (int ot = 0; ot < RG_HT; ++ot) {
        f32x4 acc = *(const f32x4*)(bias + 16 * ot + g4);
        const float* wrow = W + (16 * ot + r) * pitch + g4;
        f32x4 a[RG_HT];
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) a[it] = *(const f32x4*)(wrow + 16 * it);
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) {
          if (BF) {
            acc = rg_mfma_bf(rg_bf4(a[it]), hb[it], acc);
          } else {
            acc = rg_mfma(a[it].x, hin[it].x, acc);
            acc = rg_mfma(a[it].y, hin[it].y, acc);
            acc = rg_mfma(a[it].z, hin[it].z, acc);
            acc = rg_mfma(a[it].w, hin[it].w, acc);
          }
        }
        f32x4 y, sd;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          y[v] = act_f(1, acc[v]);
          sd[v] = act_d(1, acc[v], y[v]);
        }
        *(f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane) = sd;
        *(gp<f32x4>)(blk + m.sh[l + 1] + 16 * ot + g4) = y;
        hout[ot] = y;
      }
    } else if (nti == 1 && nto == RG_HT && act == 1 && m.act_stride && l < L - 1) {
      // the same for a first layer of 1 x 8 tiles (inputs <= 16)
#pragma unroll
      for (int ot = 0; ot < RG_HT; ++ot) {
        f32x4 acc = *(const f32x4*)(bias + 16 * ot + g4);
        const f32x4 a = *(const f32x4*)(W + (16 * ot + r) * pitch + g4);
        if (BF) {
          acc = rg_mfma_bf(rg_bf4(a), hb[0], acc);
        } else {
          acc = rg_mfma(a.x, hin[0].x, acc);
          acc = rg_mfma(a.y, hin[0].y, acc);
          acc = rg_mfma(a.z, hin[0].z, acc);
          acc = rg_mfma(a.w, hin[0].w, acc);
        }
        f32x4 y, sd;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          y[v] = act_f(1, acc[v]);
          sd[v] = act_d(1, acc[v], y[v]);
        }
        *(f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane) = sd;
        *(gp<f32x4>)(blk + m.sh[l + 1] + 16 * ot + g4) = y;
        hout[ot] = y;
      }
    } else {
#pragma unroll
    for (int ot = 0; ot < RG_HT; ++ot) {
      hout[ot] = f32x4{0.f, 0.f, 0.f, 0.f};
      if (ot < nto) {
        f32x4 acc = *(const f32x4*)(bias + 16 * ot + g4);
        const float* wrow = W + (16 * ot + r) * pitch + g4;
        // the tile's weight vectors read in one batch at constant offsets:
        // one LDS wait per tile, not per k-tile (k-tiles past nti read the
        // next rows or the zero pad after the weights, and go unused)
        f32x4 a[RG_HT];
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) a[it] = *(const f32x4*)(wrow + 16 * it);
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) {
          if (it < nti) {
            if (BF) {
              acc = rg_mfma_bf(rg_bf4(a[it]), hb[it], acc);
            } else {
              acc = rg_mfma(a[it].x, hin[it].x, acc);
              acc = rg_mfma(a[it].y, hin[it].y, acc);
              acc = rg_mfma(a[it].z, hin[it].z, acc);
              acc = rg_mfma(a[it].w, hin[it].w, acc);
            }
          }
        }
        f32x4 y, sd;
        rg_act4(act, acc, y, sd);
        if (m.act_stride) *(f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane) = sd;
        else *(gp<f32x4>)(blk + m.sd[l] + 16 * ot + g4) = sd;
        if (l < L - 1) *(gp<f32x4>)(blk + m.sh[l + 1] + 16 * ot + g4) = y;
        hout[ot] = y;
      }
    }
    }
#pragma unroll
    for (int t = 0; t < RG_HT; ++t) hin[t] = hout[t];
  }
  // backward sweep: gi = gradient entering layer l
  f32x4 gi[RG_HT], go[RG_HT];
#pragma unroll
  for (int t = 0; t < RG_HT; ++t) gi[t] = lam[t];
#pragma unroll 1
  for (int l = L - 1; l >= 0; --l) {
    const int nti = m.ti[l], nto = m.to[l], pitch = m.pitch[l];
    const float* W = lds + m.lw[l];
    if (nti == RG_HT && nto == RG_HT && m.act_stride) {
      // 8 x 8 tiles, act' in LDS: the same straight-line form as the forward
#pragma unroll
      for (int ot = 0; ot < RG_HT; ++ot) {
        const f32x4 ad = *(const f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane);
        const f32x4 d = gi[ot] * ad;
        *(gp<f32x4>)(blk + m.sd[l] + 16 * ot + g4) = d;
        gi[ot] = d;
      }
#pragma unroll
      for (int it = 0; it < RG_HT; ++it) go[it] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ot = 0; ot < RG_HT; ++ot) {
        const float* wc = W + (16 * ot + g4) * pitch + r;
        float w[RG_HT][4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int it = 0; it < RG_HT; ++it) w[it][v] = wc[v * pitch + 16 * it];
        const rg_s4 gb = rg_bf4(gi[ot]);
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) {
          if (BF) {
            go[it] = rg_mfma_bf(rg_bf4(f32x4{w[it][0], w[it][1], w[it][2], w[it][3]}), gb, go[it]);
          } else {
            go[it] = rg_mfma(w[it][0], gi[ot].x, go[it]);
            go[it] = rg_mfma(w[it][1], gi[ot].y, go[it]);
            go[it] = rg_mfma(w[it][2], gi[ot].z, go[it]);
            go[it] = rg_mfma(w[it][3], gi[ot].w, go[it]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < RG_HT; ++t) gi[t] = go[t];
      continue;
    }
    if (m.act_stride && ((nti == RG_HT && nto == 1) || (nti == 1 && nto == RG_HT))) {
      // the last layer's 8 x 1 and the first layer's 1 x 8 tiles, likewise:
      // NTO output tiles by NTI input tiles, one of them 1
      const bool last = nto == 1;
#pragma unroll
      for (int ot = 0; ot < RG_HT; ++ot) {
        if (ot > 0 && last) break;
        const f32x4 d = gi[ot] * *(const f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane);
        *(gp<f32x4>)(blk + m.sd[l] + 16 * ot + g4) = d;
        gi[ot] = d;
      }
#pragma unroll
      for (int it = 0; it < RG_HT; ++it) go[it] = f32x4{0.f, 0.f, 0.f, 0.f};
      if (last) {
        const float* wc = W + g4 * pitch + r;
        float w[RG_HT][4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int it = 0; it < RG_HT; ++it) w[it][v] = wc[v * pitch + 16 * it];
        const rg_s4 gb = rg_bf4(gi[0]);
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) {
          if (BF) {
            go[it] = rg_mfma_bf(rg_bf4(f32x4{w[it][0], w[it][1], w[it][2], w[it][3]}), gb, go[it]);
          } else {
            go[it] = rg_mfma(w[it][0], gi[0].x, go[it]);
            go[it] = rg_mfma(w[it][1], gi[0].y, go[it]);
            go[it] = rg_mfma(w[it][2], gi[0].z, go[it]);
            go[it] = rg_mfma(w[it][3], gi[0].w, go[it]);
          }
        }
      } else {
        float w[RG_HT][4];
#pragma unroll
        for (int ot = 0; ot < RG_HT; ++ot)
#pragma unroll
          for (int v = 0; v < 4; ++v) w[ot][v] = W[(16 * ot + g4 + v) * pitch + r];
#pragma unroll
        for (int ot = 0; ot < RG_HT; ++ot) {
          if (BF) {
            go[0] = rg_mfma_bf(rg_bf4(f32x4{w[ot][0], w[ot][1], w[ot][2], w[ot][3]}), rg_bf4(gi[ot]), go[0]);
          } else {
            go[0] = rg_mfma(w[ot][0], gi[ot].x, go[0]);
            go[0] = rg_mfma(w[ot][1], gi[ot].y, go[0]);
            go[0] = rg_mfma(w[ot][2], gi[ot].z, go[0]);
            go[0] = rg_mfma(w[ot][3], gi[ot].w, go[0]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < RG_HT; ++t) gi[t] = go[t];
      continue;
    }
#pragma unroll
    for (int ot = 0; ot < RG_HT; ++ot) {
      if (ot < nto) {
        gp<float> pd = blk + m.sd[l] + 16 * ot + g4;
        const f32x4 ad = m.act_stride ? *(const f32x4*)(act_lds + (m.act_off[l] + ot) * 256 + 4 * lane)
                                      : *(gp<const f32x4>)pd;
        const f32x4 d = gi[ot] * ad;
        *(gp<f32x4>)pd = d;
        gi[ot] = d;
      }
    }
    // output tile by output tile, every input tile's sum at once: the four
    // W^T row bases of an output tile serve all input tiles at constant
    // offsets (k-tiles past nti read the next rows or the zero pad, unused)
#pragma unroll
    for (int it = 0; it < RG_HT; ++it) go[it] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ot = 0; ot < RG_HT; ++ot) {
      if (ot < nto) {
        const float* wc = W + (16 * ot + g4) * pitch + r;
        float w[RG_HT][4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int it = 0; it < RG_HT; ++it) w[it][v] = wc[v * pitch + 16 * it];
        const rg_s4 gb = rg_bf4(gi[ot]);
#pragma unroll
        for (int it = 0; it < RG_HT; ++it) {
          if (it < nti) {
            if (BF) {
              go[it] = rg_mfma_bf(rg_bf4(f32x4{w[it][0], w[it][1], w[it][2], w[it][3]}), gb, go[it]);
            } else {
              go[it] = rg_mfma(w[it][0], gi[ot].x, go[it]);
              go[it] = rg_mfma(w[it][1], gi[ot].y, go[it]);
              go[it] = rg_mfma(w[it][2], gi[ot].z, go[it]);
              go[it] = rg_mfma(w[it][3], gi[ot].w, go[it]);
            }
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < RG_HT; ++t) gi[t] = go[t];
  }
#pragma unroll
  for (int t = 0; t < RG_HT; ++t) gout[t] = gi[t];
}

// One entry of the parameter part from its pre-scaled stage sums (MODE as in
// rg_mu_pass): kb = -dt sum_j b_j g~_j (MODE 0) or the g_0 sum (MODE 1), ke
// the error sum.  The error's denominator is kept at least FLT_MIN, so padded
// entries (all zero) add 0 also for atol = 0; v_rcp_f32 (1 ulp) stands in for
// the divide, whose IEEE expansion cost a dozen instructions per entry.
template <int MODE>
__device__ __forceinline__ float rg_mu_err(float kb, float ke, float atf, float rtf, float m0,
                                           float& mn, float& fl) {
  float num, mx;
  if (MODE == 0) {
    mn = m0 + kb;
    mx = fmaxf(fabsf(m0), fabsf(mn));
    num = ke;
  } else if (MODE == 1) {
    mx = fabsf(m0);
    num = kb;
  } else {
    mx = fabsf(m0);
    num = ke;
    fl = fmaxf(fl, ke != 0.f ? 1.f : 0.f);
  }
  return num * __builtin_amdgcn_rcpf(fmaxf(fmaf(mx, rtf, atf), 1.17549435e-38f));
}

// Parameter part of one trajectory's adjoint state, run by the whole wave for
// slot block `sbs` (8 stage blocks of m.sv floats; logical stage j of the
// attempt sits in block j, except j = 0 in block p0 and j = 6 in block 7-p0),
// over a row in the tile-row order of RaggedModel (mu, mu_new: Pn floats):
//  MODE 0: stage sums kb = sum_j b_j g_j, ke = sum_j btilde_j g_j with
//          g_j = -(delta_j h_j^T) (weights) and -delta_j (bias); writes
//          mu_new = mu + dt kb; returns sum (dt ke / (atol + max|mu|,|mu_new| rtol))^2
//  MODE 1: returns sum (g_0 / (atol + |mu| rtol))^2           (initial dt, d1)
//  MODE 2: returns sum ((g_1 - g_0) / (atol + |mu| rtol))^2   (initial dt, d2;
//          g_1 in block 7 - p0), *flag = 1 where g_1 != g_0
// The stage weights carry the factor of their use (-dt b_j for the update,
// dt btilde_j for the error in MODE 0, -1 in MODES 1 and 2), so an entry is
// finished by rg_mu_err alone.
// The pass streams the row: every layer's bias entries (their loads issued
// together), then the weight tiles in memory order, each tile's successor
// one row on loaded as soon as the tile is read.  It is an outlined call: inlined
// into the adjoint kernel, which runs at its register limit, its address
// registers were spilled and every load waited for the one before.
// The stage sums of a weight tile are outer products over the stages j:
//   k_b[i][o] = sum_j cb_j delta_j[o] h_j[i],  k_e likewise with ce_j,
// a GEMM with K = 7 stages (padded to 8): two v_mfma_f32_16x16x4f32 each,
// A[i][j] = h_j[i], B[j][o] = c_j delta_j[o].  Lane (r, g) supplies stage
// j = 4q + g of MFMA step q; the accumulator gives it output o = r and inputs
// i = 4g + v, which the row stores as this lane's 16-byte vector of the tile.
// Operand rows beyond the layer's inputs / outputs are zeroed once per layer,
// so a padded entry's (and a padding tile's) stage sums are exactly 0 and its
// row entry stays 0 (rows start zeroed and every write stores the whole tile).
// About one wave runs per SIMD, so the pass is bound by instruction issue:
// a layer's tile rows run along its longer tile dimension, the operand of the
// row index is picked once per row and the column index is the unrolled loop
// counter (no per-tile index arithmetic, selects or branches).
// The model struct in the kernel's argument segment (RaggedAdjArgs is the
// launch's only argument and the model its first member), read through a
// wave-uniform constant-address-space pointer: a device function gets its
// arguments in vector registers, so a struct passed to it would be read with
// per-lane FLAT loads and its layer sizes would be divergent branch
// conditions.  `kargs` is the kernel's __builtin_amdgcn_kernarg_segment_ptr().
typedef __attribute__((address_space(4))) const RaggedModel CRaggedModel;
__device__ __forceinline__ CRaggedModel& rg_uniform(uint64_t kargs) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)kargs);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(kargs >> 32));
  return *(CRaggedModel*)(((uint64_t)hi << 32) | lo);
}

template <int MODE, int RW>
__device__ __attribute__((noinline)) double rg_mu_pass(uint64_t kargs, gp<const float> sbs, int p0, double dt,
                             double atol, double rtol, gp<const float> __restrict__ mu,
                             gp<float> __restrict__ mu_new, double* flag, int lane) {
  CRaggedModel& mm = rg_uniform(kargs);
  const int r = lane & 15, g = lane >> 4;
  const float dtf = (float)dt, atf = (float)atol, rtf = (float)rtol;
  const float sb = MODE == 0 ? -dtf : -1.f, se = MODE == 0 ? dtf : -1.f;
  int boff[2];
  float cb[2], ce[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = 4 * q + g;
    int blk;
    if (MODE == 0) {
      blk = j == 0 ? p0 : (j == 6 ? 7 - p0 : (j < 7 ? j : 0));
      cb[q] = j < 6 ? cA[6][j] : 0.f;
      ce[q] = j < 7 ? cBT[j] : 0.f;
    } else if (MODE == 1) {
      blk = p0;
      cb[q] = j == 0 ? 1.f : 0.f;   // g_0
      ce[q] = 0.f;
    } else {
      blk = j == 0 ? p0 : 7 - p0;
      cb[q] = 0.f;
      ce[q] = j == 0 ? -1.f : (j == 1 ? 1.f : 0.f);  // g_1 - g_0
    }
    boff[q] = blk * mm.sv;
    cb[q] *= sb;
    ce[q] *= se;
  }
  float acc = 0.f, fl = 0.f;

  // One row buffer for the weight tiles, refilled tile by tile: as soon as a
  // tile's mu vector has been used, its register is loaded with the tile
  // RW further on (the next row, across layer ends; clamped at the last
  // tile), so a row of loads stays in flight with no copies or branches (the
  // pass runs near its register limit, and a load under a branch is waited
  // for where it joins).  The first row is requested here, before the bias
  // part, whose own loads it then overlaps.
  const gp<const f32x4> wsrc = (gp<const f32x4>)mu + lane;
  const gp<f32x4> wdst = (gp<f32x4>)mu_new + lane;
  const int tlast = mm.ntile - 1;
  f32x4 buf[RW];
#pragma unroll
  for (int k = 0; k < RW; ++k) buf[k] = wsrc[(k < tlast ? k : tlast) * 64];

  // ---- biases: lanes over the outputs o of each layer (at most two 64-wide
  // rounds: hidden widths <= 128, D <= 32).  Every load is unconditional (an
  // index out of range is clamped to a valid entry and its value dropped):
  // a load under a lane-dependent branch makes the wave wait for everything
  // in flight where the branch joins.
  const int Lm1 = mm.L - 1;
  float bm[RG_MAXL][2], bkb[RG_MAXL][2], bke[RG_MAXL][2];
#pragma unroll
  for (int l = 0; l < RG_MAXL; ++l)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int lc = l < mm.L ? l : Lm1, out = mm.out[lc];
      const int o = 64 * c + lane, oc = o < out ? o : out - 1;
      const float keep = (l < mm.L && o < out) ? 1.f : 0.f;
      const int off = mm.sd[lc] + oc;
      bm[l][c] = mu[mm.nb0[lc] + oc];
      float kb = 0.f, ke = 0.f;
      if (MODE == 0) {
#pragma unroll
        for (int j = 0; j < 7; ++j) {
          const int bj = j == 0 ? p0 : (j == 6 ? 7 - p0 : j);
          const float d = sbs[bj * mm.sv + off];
          kb += (j < 6 ? cA[6][j] : 0.f) * d;
          ke += cBT[j] * d;
        }
        kb *= sb;
        ke *= se;
      } else if (MODE == 1) {
        kb = sb * sbs[p0 * mm.sv + off];
      } else {
        ke = se * (sbs[(7 - p0) * mm.sv + off] - sbs[p0 * mm.sv + off]);
      }
      bkb[l][c] = keep * kb;
      bke[l][c] = keep * ke;
    }
#pragma unroll
  for (int l = 0; l < RG_MAXL; ++l)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int o = 64 * c + lane;
      if (l < mm.L && o < mm.out[l]) {
        float mn = 0.f;
        const float ev = rg_mu_err<MODE>(bkb[l][c], bke[l][c], atf, rtf, bm[l][c], mn, fl);
        if (MODE == 0) mu_new[mm.nb0[l] + o] = mn;
        acc += ev * ev;
      }
    }

  // ---- weight tiles: the rows of RW tiles in memory order, all layers
  // (the buffer was filled with the first row before the bias part).
  for (int l = 0; l < mm.L; ++l) {
    const int in = mm.in[l], out = mm.out[l];
    const bool ro = mm.rowout[l];
    // The lane's operands of every tile index t: h_j[16 t + r] and the
    // scaled c_j delta_j[16 t + r] of its two stages, zero beyond the layer's
    // inputs / outputs (unconditional loads of clamped indices, unwanted
    // values zeroed by a multiply -- a select here compiles to branches, see
    // the bias part; every clamped index reads finite stage data).  They are
    // sorted into the operands of the row index (tB, tE) and of the column
    // index (cB, cE) of the update and error sums; rows along the inputs
    // (ro = 0) store their tiles transposed, so that both orders run the same
    // code: the update sum is MFMA(row operand, column operand).
    float tB[RW][2], tE[RW][2], cB[RW][2], cE[RW][2];
#pragma unroll
    for (int t = 0; t < RW; ++t) {
      const int i = 16 * t + r;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const float hv = sbs[boff[q] + mm.sh[l] + (i < in ? i : in - 1)];
        const float dv = sbs[boff[q] + mm.sd[l] + (i < out ? i : out - 1)];
        const float h = (i < in ? 1.f : 0.f) * hv;
        const float dm = (i < out ? 1.f : 0.f) * dv;
        const float bq = cb[q] * dm, eq = ce[q] * dm;
        tB[t][q] = ro ? h : bq;
        tE[t][q] = ro ? h : eq;
        cB[t][q] = ro ? bq : h;
        cE[t][q] = ro ? eq : h;
      }
    }
    const int T0 = (int)(mm.nw0[l] >> 8), nr = mm.nrow[l];
    for (int a = 0; a < nr; ++a) {
      float rb0 = 0.f, rb1 = 0.f, re0 = 0.f, re1 = 0.f;
#pragma unroll
      for (int t = 0; t < RW; ++t)
        if (t == a) { rb0 = tB[t][0]; rb1 = tB[t][1]; re0 = tE[t][0]; re1 = tE[t][1]; }
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        // tile T: stage sums, entries, store; then the refill
        const int T = T0 + a * RW + k;
        f32x4 kb = {0.f, 0.f, 0.f, 0.f}, ke = kb, mn = kb;
        if (MODE != 2) {
          kb = rg_mfma(rb0, cB[k][0], kb);
          kb = rg_mfma(rb1, cB[k][1], kb);
        }
        if (MODE != 1) {
          ke = rg_mfma(re0, cE[k][0], ke);
          ke = rg_mfma(re1, cE[k][1], ke);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float mv = 0.f;
          const float ev = rg_mu_err<MODE>(kb[v], ke[v], atf, rtf, buf[k][v], mv, fl);
          mn[v] = mv;
          acc = fmaf(ev, ev, acc);
        }
        if (MODE == 0) wdst[T * 64] = mn;
        const int Tn = T + RW;
        buf[k] = wsrc[(Tn < tlast ? Tn : tlast) * 64];
      }
    }
  }

  double sum = (double)acc, flg = (double)fl;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off);
    flg = fmax(flg, __shfl_xor(flg, off));
  }
  if (flag) *flag = flg;
  return sum;
}

// The pass for the model's row width (1, 2, 4 or 8 tiles) and mode.
__device__ __forceinline__ double rg_mu_call(uint64_t kargs, int rw, int mode, gp<const float> sbs,
                                             int p0, double dt, double atol, double rtol,
                                             gp<const float> mu, gp<float> mun, double* fl, int lane) {
#define RG_MU_RW(R)                                                                          \
  if (rw == R) {                                                                            \
    if (mode == 0) return rg_mu_pass<0, R>(kargs, sbs, p0, dt, atol, rtol, mu, mun, nullptr, lane); \
    if (mode == 1) return rg_mu_pass<1, R>(kargs, sbs, p0, dt, atol, rtol, mu, nullptr, nullptr, lane); \
    return rg_mu_pass<2, R>(kargs, sbs, p0, dt, atol, rtol, mu, nullptr, fl, lane);                \
  }
  RG_MU_RW(1)
  RG_MU_RW(2)
  RG_MU_RW(4)
#undef RG_MU_RW
  if (mode == 0) return rg_mu_pass<0, RG_HT>(kargs, sbs, p0, dt, atol, rtol, mu, mun, nullptr, lane);
  if (mode == 1) return rg_mu_pass<1, RG_HT>(kargs, sbs, p0, dt, atol, rtol, mu, nullptr, nullptr, lane);
  return rg_mu_pass<2, RG_HT>(kargs, sbs, p0, dt, atol, rtol, mu, nullptr, fl, lane);
}

// A finished trajectory's row from the tile-row order to the flat
// parameter order (its first P floats); `tmp` holds the tile-major row.
__device__ void rg_row_to_flat(const RaggedModel& m, gp<const float> __restrict__ tmp,
                               gp<float> __restrict__ row, int lane) {
  const int r = lane & 15, g = lane >> 4;
  for (int l = 0; l < m.L; ++l) {
    const int in = m.in[l], out = m.out[l], nti = m.ti[l], nto = m.to[l];
    const gp<const f32x4> src = (gp<const f32x4>)(tmp + m.nw0[l]) + lane;
    for (int t = 0; t < m.nrow[l] * m.rw; ++t) {
      const int a = t / m.rw, k = t - a * m.rw;
      const int it = m.rowout[l] ? a : k, ot = m.rowout[l] ? k : a;
      if (it >= nti || ot >= nto) continue;   // padding tile
      const f32x4 v4 = src[t * 64];
#pragma unroll
      for (int v = 0; v < 4; ++v) {   // (tiles of rows along the inputs are transposed)
        const int o = 16 * ot + (m.rowout[l] ? r : 4 * g + v);
        const int i = 16 * it + (m.rowout[l] ? 4 * g + v : r);
        if (o < out && i < in) row[m.w0[l] + (int64_t)i * out + o] = v4[v];
      }
    }
    for (int o = lane; o < out; o += 64) row[m.b0[l] + o] = tmp[m.nb0[l] + o];
  }
}

// u(tau) of the slot's trajectory from its forward tape: the stored state when
// tau is a step time, else the step's free interpolant (dt = t_{j+1} - t_j);
// (cj, tcj, tnx): a record at or before tau's step and the time after it
template <int DT>
__device__ __forceinline__ void rg_tape_u(const RaggedAdjArgs& a, int cj, double tcj, double tnx,
                                          double tau, f32x4 (&uu)[DT], int lane) {
  const int g4 = 4 * (lane >> 4);
  int j = cj;
  double tj = tcj, tn = tnx;
  const gp<const int> rprev = glb(a.rec_prev);
  const gp<const double> rt = glb(a.rec_t);
  while (tj > tau && rprev[j] >= 0) {
    tn = tj;
    j = rprev[j];
    tj = rt[j];
  }
  const gp<const float> rd = glb(a.rec_data) + (int64_t)j * 8 * a.Dp;
  if (tj == tau || !(tn > tj)) {
#pragma unroll
    for (int t = 0; t < DT; ++t) uu[t] = *(gp<const f32x4>)(rd + 16 * t + g4);
    return;
  }
  const double dts = tn - tj;
  float w[7];
  interp_weights((tau - tj) / dts, w);
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    f32x4 sm = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 7; ++i) sm += w[i] * *(gp<const f32x4>)(rd + (1 + i) * a.Dp + 16 * t + g4);
    uu[t] = *(gp<const f32x4>)(rd + 16 * t + g4) + (float)dts * sm;
  }
}

template <int DT, bool BF>
__global__ void __launch_bounds__(256) k_ragged_adj(RaggedAdjArgs a) {
  static_assert(offsetof(RaggedAdjArgs, m) == 0, "the model leads the argument struct");
  extern __shared__ __attribute__((aligned(16))) float rg_lds[];
  const RaggedModel& m = a.m;
  const uint64_t kargs = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
  rg_load_weights(m, rg_lds);
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const int64_t wslot0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * RG_SLOTS;
  // stage blocks of the slot's VJPs; the columns past spw never hold a
  // trajectory, and their VJP writes (garbage that stays in its own MFMA
  // column) go to one sink block shared by all of them, which stays in L2
  const gp<float> sb = r < a.spw ? glb(a.sbuf) + (wslot0 + r) * 8 * m.sv
                                 : glb(a.sbuf) + a.nslots_all * 8 * m.sv;
  const int D = a.D;
  const int64_t P = m.P;
  const double nel = (double)D + (double)P;
  const float atol = (float)a.abstol, rtol = (float)a.reltol;

  int traj = -1, phase = 0, stage = 0, p0 = 0, pad = 0, fsal_re = 0, muflip = 0, cj = 0;
  bool need_start = true, first_start = true;
  SlotCtl c{};
  f32x4 lam[DT], lnew[DT], kl[7][DT], uc[DT];
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    lam[t] = lnew[t] = uc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 7; ++j) kl[j][t] = lam[t];
  }
  int64_t s_lo = 0, snext = -1, nf = 0, nacc = 0, nrej = 0;
  double tcj = 0.0, tnx = 0.0, dt0 = 0.0, d0 = 0.0, d1 = 0.0, s1l = 0.0;

  for (;;) {
    if (need_start) {
      need_start = false;
      // the slot's own position first (slots >= spw stay idle), then
      // positions from the queue
      int64_t pos = r < a.spw ? wslot0 / RG_SLOTS * a.spw + r : a.norder;
      if (!first_start) {
        int q = 0;
        if (g == 0) q = atomicAdd(a.queue, 1);
        pos = a.nslots + rg_bcast(q, lane);
      }
      first_start = false;
      traj = pos < a.norder ? a.order[pos] : -1;
      if (traj >= 0) {
        c = SlotCtl{};
        c.t = a.t1[traj];
        c.t_end = a.t0;
        c.tdir = -1.0;
        c.dtmax_abs = a.dtmax > 0 ? a.dtmax : fabs(c.t - c.t_end);
        c.dtmin = a.dtmin > 0 ? a.dtmin : fmax(dspacing(c.t), dspacing(c.t_end));
        c.maxiters = a.maxiters;
        c.adaptive = a.adaptive;
        c.dt_fixed = a.dt_init;
        c.qold = QOLDINIT;
        c.q11 = 1.0;
        s_lo = a.save_off[traj];
        snext = a.save_off[traj + 1] - 1;
        int64_t jump = -1;
        if (snext >= s_lo && a.save_t[snext] == c.t) { jump = snext; --snext; }
        double s0 = 0.0;
#pragma unroll
        for (int t = 0; t < DT; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int f = 16 * t + 4 * g + v;
            const float l = (jump >= 0 && f < D) ? a.dLdout[jump * D + f] : 0.f;
            lam[t][v] = l;
            if (f < D) {
              const double q = (double)l / (a.abstol + fabs((double)l) * a.reltol);
              s0 += q * q;
            }
          }
        s0 = rg_slot_sum(s0);
        d0 = sqrt(s0 / nel);
        cj = a.last_rec[traj];
        tcj = a.rec_t[cj];
        tnx = INFINITY;
        p0 = 0;
        muflip = 0;
        fsal_re = 0;
        phase = 0;
        stage = 0;
        nf = nacc = nrej = 0;
      }
    }
    if (!__any(traj >= 0)) break;
    // ---- VJP input of this round: stage time and adjoint state
    double tau = c.t;
    f32x4 ls[DT];
    const float h = (float)c.dt;
    int blkid = p0;
#pragma unroll
    for (int t = 0; t < DT; ++t) ls[t] = lam[t];
    if (traj >= 0) {
      if (phase == 1) {
        tau = c.t + c.tdir * dt0;
        blkid = 7 - p0;
#pragma unroll
        for (int t = 0; t < DT; ++t) ls[t] = lam[t] + (float)(c.tdir * dt0) * kl[0][t];
      } else if (phase == 2) {
        tau = c.t + dC[stage] * c.dt;
        blkid = stage == 0 ? p0 : (stage == 6 ? 7 - p0 : stage);
        if (stage > 0) {
#pragma unroll
          for (int t = 0; t < DT; ++t) ls[t] = lam[t] + h * rg_stage_sum(kl, stage, t);
        }
      }
      rg_tape_u<DT>(a, cj, tcj, tnx, tau, uc, lane);
    }
    f32x4 x0[RG_HT], lg[RG_HT], gv[RG_HT];
#pragma unroll
    for (int t = 0; t < RG_HT; ++t) {
      x0[t] = lg[t] = f32x4{0.f, 0.f, 0.f, 0.f};
      if (t < DT && traj >= 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) x0[t][q] = rg_map(m.map, uc[t][q]);
        lg[t] = ls[t];
      }
    }
    rg_vjp<BF>(m, rg_lds, rg_lds + m.lds_floats + (threadIdx.x >> 6) * m.act_stride, x0, lg,
           sb + blkid * m.sv, gv, lane);
    // k = -(lambda^T df/du): input-map derivative 3u^2 for x.^3
    f32x4 kn[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      f32x4 v = gv[t];
      if (m.map == 1) v *= 3.f * uc[t] * uc[t];
      kn[t] = -v;
    }
    // ---- consume, part 1 (adjoint-state terms); mu-pass requests
    int mode = -1;   // mu pass: 0 step, 1 d1, 2 d2
    double sl = 0.0, fll = 0.0;
    if (traj >= 0) {
      ++nf;
      if (phase == 0) {
#pragma unroll
        for (int t = 0; t < DT; ++t) {
          kl[0][t] = kn[t];
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (16 * t + 4 * g + v < D) {
              const float sk = atol + fabsf(lam[t][v]) * rtol;
              const double q = (double)(kn[t][v] / sk);
              sl += q * q;
            }
        }
        mode = 1;
      } else if (phase == 1) {
#pragma unroll
        for (int t = 0; t < DT; ++t) {
          kl[6][t] = kn[t];
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (16 * t + 4 * g + v < D) {
              const float sk = atol + fabsf(lam[t][v]) * rtol;
              const double q = (double)((kn[t][v] - kl[0][t][v]) / sk);
              sl += q * q;
              if (kn[t][v] != kl[0][t][v]) fll = 1.0;
            }
        }
        mode = 2;
      } else {
#pragma unroll
        for (int j = 0; j < 7; ++j)
          if (j == stage)
#pragma unroll
            for (int t = 0; t < DT; ++t) kl[j][t] = kn[t];
        if (stage == 6) {
#pragma unroll
          for (int t = 0; t < DT; ++t) {
            lnew[t] = ls[t];
            f32x4 ut = cBT[6] * kl[6][t];
#pragma unroll
            for (int j = 0; j < 6; ++j) ut += cBT[j] * kl[j][t];
            ut *= h;
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (16 * t + 4 * g + v < D) {
                const float sk = atol + fmaxf(fabsf(lam[t][v]), fabsf(lnew[t][v])) * rtol;
                const double q = (double)(ut[v] / sk);
                sl += q * q;
                if (isnan(lnew[t][v])) fll = 1.0;
              }
          }
          mode = 0;
        }
      }
      sl = rg_slot_sum(sl);
      fll = rg_slot_sum(fll);
    }
    // ---- parameter part, slot by slot with the whole wave
    __threadfence_block();
    double smu = 0.0, flmu = 0.0;
    for (int s = 0; s < RG_SLOTS; ++s) {
      const int ms = __shfl(mode, s);
      if (ms < 0) continue;
      const int trs = __shfl(traj, s);
      const int p0s = __shfl(p0, s);
      const int mfs = __shfl(muflip, s);
      const double dts = __shfl(c.dt, s);
      const gp<const float> sbs = glb(a.sbuf) + (wslot0 + s) * 8 * m.sv;
      const gp<float> row = glb(a.dLdp_traj) + (int64_t)trs * m.Pn;
      const gp<float> tmp = glb(a.mu_tmp) + (wslot0 + s) * m.Pn;
      const gp<const float> mu = mfs ? tmp : row;
      const gp<float> mun = mfs ? row : tmp;
      double fl = 0.0, sm;
      sm = rg_mu_call(kargs, m.rw, ms, sbs, p0s, dts, a.abstol, a.reltol, mu, mun, &fl, lane);
      if (r == s) { smu = sm; flmu = fl; }
    }
    // ---- consume, part 2: control
    bool fin = false;
    if (traj < 0) {
      // idle slot: nothing to decide
    } else if (phase == 0) {
      s1l = sl;
      d1 = sqrt((sl + smu) / nel);
      if (a.dt_init > 0.0) {
        c.dt = c.tdir * a.dt_init;
        c.dtpropose = c.dt;
        c.last_accept = 0;
        loopheader(&c, snext >= s_lo ? a.save_t[snext] : c.t_end);
        phase = 2;
        stage = 1;
        fin = c.done;
      } else {
        const double dtmin = nextafter(fmax(c.dtmin, dspacing(c.t)), INFINITY);
        const double smalldt = fmax(dtmin, 1e-6);
        dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
        dt0 = fmin(dt0, c.dtmax_abs);
        pad = 0;
        if (!(dt0 >= 10 * 2.220446049250313e-16)) { dt0 = smalldt; pad = 1; }
        phase = 1;
      }
    } else if (phase == 1) {
      const double dtmin = nextafter(fmax(c.dtmin, dspacing(c.t)), INFINITY);
      if (pad) {
        c.dt = c.tdir * dt0;
      } else if (fmax(fll, flmu) == 0.0) {
        c.dt = c.tdir * fmax(dtmin, 100 * dt0);
      } else {
        const double d2 = sqrt((sl + smu) / nel) / dt0;
        const double mx = fmax(d1, d2);
        const double dt1 = mx <= 1e-15 ? fmax(1e-6, dt0 * 1e-3) : pow(10.0, -(2.0 + log10(mx)) / 5.0);
        c.dt = c.tdir * fmax(dtmin, fmin(fmin(100 * dt0, dt1), c.dtmax_abs));
      }
      c.dtpropose = c.dt;
      c.last_accept = 0;
      loopheader(&c, snext >= s_lo ? a.save_t[snext] : c.t_end);
      phase = 2;
      stage = 1;
      fin = c.done;
    } else if (stage < 6) {
      ++stage;
    } else {
      const double eest = c.adaptive ? sqrt((sl + smu) / nel) : 0.0;
      fsal_re = 0;
      if (c.adaptive && isnan(eest)) {
        c.retcode = 3;
        c.done = 1;
      } else if (pi_control(&c, eest)) {
        ++nacc;
        const double tstop = snext >= s_lo ? a.save_t[snext] : c.t_end;
        c.t = snap_t(c.t, c.t + c.dt, tstop);
        muflip ^= 1;
        p0 = 7 - p0;  // FSAL: the k7 block is the next attempt's k1 block
#pragma unroll
        for (int t = 0; t < DT; ++t) { lam[t] = lnew[t]; kl[0][t] = kl[6][t]; }
        if (snext >= s_lo && c.t == a.save_t[snext]) {
#pragma unroll
          for (int t = 0; t < DT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int f = 16 * t + 4 * g + v;
              if (f < D) lam[t][v] += a.dLdout[snext * D + f];
            }
          --snext;
          fsal_re = 1;
        }
        // tape cursor back to the record holding the new time
        while (tcj > c.t && a.rec_prev[cj] >= 0) {
          tnx = tcj;
          cj = a.rec_prev[cj];
          tcj = a.rec_t[cj];
        }
        if (!c.adaptive) c.dtpropose = c.dt;
        c.last_accept = 1;
        if (fll != 0.0) { c.retcode = 6; c.done = 1; }
        else if (c.t == c.t_end) { c.retcode = 1; c.done = 1; }
        else c.dt = c.dtpropose;
      } else {
        ++nrej;
        c.last_accept = 0;
        c.dt = c.dt / fmin(1.0 / QMIN, c.q11 / GAMMA);
      }
      if (!c.done) loopheader(&c, snext >= s_lo ? a.save_t[snext] : c.t_end);
      stage = fsal_re ? 0 : 1;
      fin = c.done;
    }
    if (fin) {  // (only slots with a trajectory reach fin)
#pragma unroll
      for (int t = 0; t < DT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int f = 16 * t + 4 * g + v;
          if (f < D) a.dLdu0[(int64_t)traj * D + f] = lam[t][v];
        }
      if (g == 0) {
        a.retcode[traj] = c.retcode;
        int* ts = a.traj_stats + 4 * (int64_t)traj;
        ts[0] = (int)nf; ts[1] = (int)nacc; ts[2] = (int)nrej; ts[3] = 0;
        atomicAdd(a.stats + 0, (unsigned long long)nf);
        atomicAdd(a.stats + 1, (unsigned long long)nacc);
        atomicAdd(a.stats + 2, (unsigned long long)nrej);
      }
      need_start = true;
    }
    // the finished trajectory's parameter gradient ends in its own row in the
    // flat parameter order (whole wave, converged): the tile-major row is
    // first copied to the slot's row when it is not there already
    if (__any(fin)) {
      for (int s = 0; s < RG_SLOTS; ++s) {
        if (!__shfl((int)fin, s)) continue;
        const int trs = __shfl(traj, s);
        const gp<float> tmp = glb(a.mu_tmp) + (wslot0 + s) * m.Pn;
        const gp<float> row = glb(a.dLdp_traj) + (int64_t)trs * m.Pn;
        if (!__shfl(muflip, s)) {
          for (int64_t i = 4 * lane; i < m.Pn; i += 256) *(gp<f32x4>)(tmp + i) = *(gp<const f32x4>)(row + i);
          __threadfence_block();
        }
        rg_row_to_flat(m, tmp, row, lane);
      }
    }
  }
}

// `waves` (1..4) waves of RG_SLOTS slots per workgroup
void launch_ragged_adj(const RaggedAdjArgs& a, int grid, int waves, hipStream_t s) {
  const size_t lds = ((size_t)a.m.lds_floats + (size_t)waves * a.m.act_stride) * sizeof(float);
  auto go = [&](auto kern) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * waves), lds, s, a);
  };
  if (a.Dp <= 16) {
    if (a.vjp_bf16) go(k_ragged_adj<1, true>);
    else go(k_ragged_adj<1, false>);
  } else {
    if (a.vjp_bf16) go(k_ragged_adj<2, true>);
    else go(k_ragged_adj<2, false>);
  }
}

// dL/dp = sum over trajectories of their FP32 rows (first P of every `stride`
// floats) in FP64, in a fixed order: RG_SUM_CHUNKS consecutive ranges of
// trajectories, each summed in trajectory order (blockIdx.y: the range), then
// the ranges' sums in range order.  One thread per parameter and range keeps
// thousands of blocks streaming the rows (a thread per parameter over all
// trajectories left most CUs idle and every thread one long dependent chain).
__global__ void k_ragged_sum_part(const float* rows, int64_t n, int64_t P, int64_t stride,
                                  double* part) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int64_t c = blockIdx.y, k0 = n * c / RG_SUM_CHUNKS, k1 = n * (c + 1) / RG_SUM_CHUNKS;
  double s = 0.0;
  for (int64_t k = k0; k < k1; ++k) s += (double)rows[k * stride + i];
  part[c * P + i] = s;
}
__global__ void k_ragged_sum_fold(const double* part, int64_t P, double* out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= P) return;
  double s = 0.0;
  for (int c = 0; c < RG_SUM_CHUNKS; ++c) s += part[c * P + i];
  out[i] = s;
}
void launch_ragged_sum(const float* rows, int64_t n, int64_t P, int64_t stride, double* part,
                       double* out, hipStream_t s) {
  const int64_t nb = (P + 255) / 256;
  hipLaunchKernelGGL(k_ragged_sum_part, dim3((unsigned)nb, RG_SUM_CHUNKS), dim3(256), 0, s, rows, n,
                     P, stride, part);
  hipLaunchKernelGGL(k_ragged_sum_fold, dim3((unsigned)nb), dim3(256), 0, s, part, P, out);
}
