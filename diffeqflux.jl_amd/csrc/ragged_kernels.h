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
// share the machine until the queue is empty.  Every trajectory's arithmetic
// depends only on its own column, so results do not depend on scheduling.

constexpr int RG_SLOTS = 16;

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct SlotCtl {
  double t, dt, tprev, dtpropose, qold, q11, eest, tdir, dtmax_abs, dtmin, t_end, dt_fixed;
  int64_t iter, maxiters;
  int adaptive, retcode, done, last_accept;
};

__device__ __forceinline__ f32x4 rg_mfma(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
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
// map applied) -> f (DT tiles).
template <int DT>
__device__ __forceinline__ void rg_mlp(const RaggedModel& m, const float* lds,
                                       const f32x4 (&x)[DT], f32x4 (&f)[DT], int lane) {
  const int L = m.L;
  if (L == 1) {
    rg_layer<DT, DT>(lds + m.lw[0], m.pitch[0], lds + m.lb[0], m.act[0], x, m.ti[0], m.to[0], f, lane);
    return;
  }
  f32x4 ha[RG_HT], hb[RG_HT];
  rg_layer<DT, RG_HT>(lds + m.lw[0], m.pitch[0], lds + m.lb[0], m.act[0], x, m.ti[0], m.to[0], ha, lane);
#pragma unroll
  for (int l = 1; l < RG_MAXL - 1; ++l) {
    if (l < L - 1) {
      if (l & 1)
        rg_layer<RG_HT, RG_HT>(lds + m.lw[l], m.pitch[l], lds + m.lb[l], m.act[l], ha, m.ti[l], m.to[l], hb, lane);
      else
        rg_layer<RG_HT, RG_HT>(lds + m.lw[l], m.pitch[l], lds + m.lb[l], m.act[l], hb, m.ti[l], m.to[l], ha, lane);
    }
  }
  const int l = L - 1;
  if (l & 1) rg_layer<RG_HT, DT>(lds + m.lw[l], m.pitch[l], lds + m.lb[l], m.act[l], ha, m.ti[l], m.to[l], f, lane);
  else rg_layer<RG_HT, DT>(lds + m.lw[l], m.pitch[l], lds + m.lb[l], m.act[l], hb, m.ti[l], m.to[l], f, lane);
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
