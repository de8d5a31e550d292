// ut_kernels.hip -- gfx950 kernels of the batched MCMC path around the per-chain
// kernels of ut_chain.h: the pooled mass replay, Philox fills, and masked NUTS (the
// persistent transition as a resumable per-leaf state machine).
#include "ut_chain.h"

namespace ut {

// Pooled running-variance replay (reference warmup.jl:262-327 driving
// nuts/adaptation.jl:30-41,170-260 and the clip constants of data_types.jl:189-193).
// The reference folds each warmup iteration's P x C positions into one clipped,
// weighted Welford state, chain by chain.  The state's scalars (count, weight sums)
// evolve identically for every parameter, so each thread replays the same sequence
// for its own parameters; no synchronisation is needed.
constexpr int REPLAY_THREADS = 256;
constexpr int REPLAY_PER_THREAD = 4;

__device__ inline double window_progress(long long count, int wl) {
  if (wl <= 4) return 1.0;
  const double num = (double)(count - 4 > 0 ? count - 4 : 0);
  const double r = num / (double)(wl - 4);
  return r < 1.0 ? r : 1.0;
}

// Each iteration's history is staged through LDS in blocks of `cb` chains (positions
// as one contiguous copy, weights computed lane-parallel), so the sequential fold reads
// LDS instead of waiting on a global load per step.  RPT parameters per thread (1 when
// P <= REPLAY_THREADS); the fold is latency-bound (one binary64 dependency chain per
// parameter), so it avoids branches: the running weight sums are a branch-free loop,
// the next chain's staged values are read while the current one is folded, and the
// clip's divide and square root run only when some lane of the wavefront needs them.
template <int RPT>
__global__ __launch_bounds__(REPLAY_THREADS) void k_mass_replay(
    const float* __restrict__ hist_pos, const double* __restrict__ hist_acc,
    const uint8_t* __restrict__ hist_flags, int C, int P, int it_lo, int it_hi, int wl,
    int use_accepted, double reg, int min_samples, double* im_d, float* im_f, double* summary, int cb) {
  extern __shared__ double replay_sh[];
  double* sw = replay_sh;                  // cb weights
  double* sfrac = sw + cb;                 // cb w / (running weight sum)
  double* sclip = sfrac + cb;              // cb clip scale, or -1 before the clip starts
  double* sden = sclip + cb;               // cb max(count - 1, 1)
  long long* scnt = (long long*)(sden + cb);  // cb running count before the step
  double* sstate = (double*)(scnt + cb);   // running weight sum / square sum / count after the block
  float* sx = (float*)(sstate + 4);        // cb x P positions
  const int t = threadIdx.x;
  double mean[RPT], m2[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) { mean[j] = 0.0; m2[j] = 0.0; }
  long long count = 0;
  double wsum = 0.0, wsq = 0.0;
  for (int it = it_lo; it <= it_hi; ++it) {
    const double progress = window_progress(count, wl);
    for (int c0 = 0; c0 < C; c0 += cb) {
      const int nb = C - c0 < cb ? C - c0 : cb;
      const long long row0 = (long long)(it - 1) * C + c0;
      __syncthreads();
      {
        // SU loads in flight per thread (indices clamped into the block), then the LDS stores
        constexpr int SU = 16;
        const int nbp = nb * P;
        const float* src = hist_pos + row0 * P;
        for (int e0 = t; e0 < nbp; e0 += (int)blockDim.x * SU) {
          float v[SU];
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            const int e = e0 + u * (int)blockDim.x;
            v[u] = src[e < nbp ? e : nbp - 1];
          }
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            const int e = e0 + u * (int)blockDim.x;
            if (e < nbp) sx[e] = v[u];
          }
        }
      }
      for (int c = t; c < nb; c += blockDim.x) {
        const uint8_t fl = hist_flags[row0 + c];
        const double ap = hist_acc[row0 + c];
        double w;
        if (fl & FLAG_DIVERGENT) w = 0.0;
        else if (use_accepted && (fl & FLAG_ACCEPTED)) w = 1.0;
        else if (!isfinite_d(ap)) w = 0.0;
        else {
          const double rej = ap < 0.0 ? 0.0 : (ap > 1.0 ? 1.0 : ap);
          w = 1.0 + (rej - 1.0) * progress;
        }
        sw[c] = w;
      }
      __syncthreads();
      // the state's scalars in chain order: running count and weight sums, by the first
      // wavefront in groups of 64 chains: lane k holds chain c0 + k's weight, every lane
      // runs the same sums in chain order on the weights read across lanes (the weight
      // sum's additions keep their order; a zero weight leaves the sums unchanged), and
      // lane k keeps the state before and after chain k's step ...
      if (t < WAVE) {
        long long cnt = count;
        double ws = wsum, wq = wsq;
        for (int g0 = 0; g0 < nb; g0 += WAVE) {
          const int c = g0 + t;
          const double wl = sw[c < nb ? c : nb - 1];
          const int m = nb - g0 < WAVE ? nb - g0 : WAVE;
          long long my_cnt = 0;
          double my_ws = 0.0;
          for (int k = 0; k < m; ++k) {
            const double w = readlane(wl, k);
            my_cnt = t == k ? cnt : my_cnt;
            const bool on = w != 0.0;
            ws = on ? ws + w : ws;
            wq = on ? wq + w * w : wq;
            cnt += on ? 1 : 0;
            my_ws = t == k ? ws : my_ws;
          }
          if (c < nb) {
            scnt[c] = my_cnt;
            sfrac[c] = my_ws;
          }
        }
        if (t == 0) {
          sstate[0] = ws;
          sstate[1] = wq;
          sstate[2] = (double)cnt;
        }
      }
      __syncthreads();
      // ... then every step's clip scale, divisor and weight fraction in parallel
      for (int c = t; c < nb; c += blockDim.x) {
        const double w = sw[c];
        if (w == 0.0) continue;
        const long long cnt = scnt[c];
        double clip_scale = -1.0;
        if (cnt >= 4) clip_scale = cnt > 4 ? 8.0 + (5.0 - 8.0) * window_progress(cnt, wl) : 8.0;
        sclip[c] = clip_scale;
        sden[c] = (double)(cnt - 1 > 1 ? cnt - 1 : 1);
        sfrac[c] = w / sfrac[c];
      }
      __syncthreads();
      // software-pipelined: chain c + 1's staged values are read from LDS while chain c
      // is folded, so the fold's dependency chain does not wait on LDS (the entries of
      // zero-weight chains are read but not used; parameters >= P read position 0 and
      // fold zeros, which no result uses)
      double w_n = sw[0], cs_n = sclip[0], den_n = sden[0], fr_n = sfrac[0];
      float x_n[RPT];
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int p = t + j * REPLAY_THREADS;
        const float v = sx[p < P ? p : 0];
        x_n[j] = p < P ? v : 0.0f;
      }
      for (int c = 0; c < nb; ++c) {
        const double w = w_n;
        const double clip_scale = cs_n;
        const double denom_c = den_n;
        const double frac = fr_n;
        double x[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) x[j] = (double)x_n[j];
        {
          const int cn = c + 1 < nb ? c + 1 : c;
          w_n = sw[cn];
          cs_n = sclip[cn];
          den_n = sden[cn];
          fr_n = sfrac[cn];
#pragma unroll
          for (int j = 0; j < RPT; ++j) {
            const int p = t + j * REPLAY_THREADS;
            const float v = sx[cn * P + (p < P ? p : 0)];
            x_n[j] = p < P ? v : 0.0f;
          }
        }
        if (w == 0.0) continue;
        // The step is computed for the unclipped deviation first, alongside the clip
        // test, so that the test's latency overlaps the update; the rare steps whose
        // deviation lies beyond the bound are then recomputed from the clipped one.
        // |delta| <= bound  <=>  delta^2 denom_c <= clip_scale^2 max(m2, 1e-3 denom_c); with
        // a 1e-12 relative margin the test below implies the exact comparison, so the
        // divide and square root are only evaluated for samples near or beyond the bound.
        const bool clip = clip_scale >= 0.0;
        const double inside = clip_scale * clip_scale * (1.0 - 1e-12);
        const double floor_m2 = 1e-3 * denom_c;
        double dl[RPT], mn[RPT], mm[RPT];
        bool in[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) dl[j] = 0.0;
        bool out = false;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          double xf = x[j];
          if (clip) {
            dl[j] = x[j] - mean[j];
            xf = mean[j] + dl[j];
          }
          const double delta = xf - mean[j];
          mn[j] = mean[j] + frac * delta;
          const double delta2 = xf - mn[j];
          mm[j] = m2[j] + (w * delta) * delta2;
          in[j] = !clip || dl[j] * dl[j] * denom_c <= inside * (m2[j] > floor_m2 ? m2[j] : floor_m2);
          out = out || !in[j];
        }
        if (wave_any(out)) {
#pragma unroll
          for (int j = 0; j < RPT; ++j) {
            if (!in[j]) {
              const double var = m2[j] / denom_c;
              const double bound = clip_scale * sqrt_d(var > 1e-3 ? var : 1e-3);
              const double d = dl[j] > bound ? bound : (dl[j] < -bound ? -bound : dl[j]);
              const double xf = mean[j] + d;
              const double delta = xf - mean[j];
              mn[j] = mean[j] + frac * delta;
              const double delta2 = xf - mn[j];
              mm[j] = m2[j] + (w * delta) * delta2;
            }
          }
        }
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          mean[j] = mn[j];
          m2[j] = mm[j];
        }
      }
      wsum = sstate[0];
      wsq = sstate[1];
      count = (long long)sstate[2];
    }
  }
  const double eff = wsq <= 0.0 ? 0.0 : (wsum * wsum) / wsq;
  const bool update = eff >= (double)min_samples;
  if (update) {
    const double denom = wsum - wsq / wsum;
    const double shrink = eff / (eff + 5.0);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int p = t + j * REPLAY_THREADS;
      if (p < P) {
        double v;
        if (eff < 2.0 || denom <= 0.0) v = 1.0;
        else {
          v = shrink * (m2[j] / denom) + (1.0 - shrink);
          v = v > reg ? v : reg;
        }
        im_d[p] = v;
        im_f[p] = (float)v;
      }
    }
  }
  if (t == 0) {
    summary[0] = (double)count;
    summary[1] = wsum;
    summary[2] = wsq;
    summary[3] = eff;
    summary[4] = update ? 1.0 : 0.0;
  }
}

__global__ void k_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream_id,
                              uint32_t first, long long n, uint32_t* blocks, float* uf, float* nf,
                              double* ud) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const philox_block b = draw_block(chain, iteration, stream_id, first + (uint32_t)i);
  if (blocks) {
    blocks[4 * i + 0] = b.w[0];
    blocks[4 * i + 1] = b.w[1];
    blocks[4 * i + 2] = b.w[2];
    blocks[4 * i + 3] = b.w[3];
  }
  if (uf) uf[i] = u01_f(b.w[0]);
  if (nf) nf[i] = box_muller_f(b);
  if (ud) ud[i] = u01_d(b.w[0], b.w[1]);
}

// ------------------------------------------------------------------ masked NUTS
// The persistent transition above restated as a resumable state machine: the same
// arithmetic in the same order (so a wavefront-per-chain model gives the same draws
// bit for bit), with the chain's tree state kept in HBM between leaves and each
// leaf's gradient taken from one batched call for all chains (the matrix-core GLM,
// the GP, a flow-wrapped model, or the wavefront models' density kernel).  A chain
// that finishes an iteration starts its next one in the same step, so chains are
// never idle inside a launch range; chains past the range's end are masked.
// Plans without a wavefront model use the layout of this identity model.
template <int KK>
struct IdentityLayout {
  static constexpr int K = KK;
  int P;
  __device__ int api(int lane, int slot) const {
    const int i = slot * WAVE + lane;
    return i < P ? i : -1;
  }
};

enum { MN_START = 0, MN_FIRST = 1, MN_LEAF = 2, MN_DONE = 3 };
enum { MN_P = 0, MN_Q, MN_G, MN_Q0, MN_G0, MN_LQ, MN_LP, MN_LG, MN_RQ, MN_RP, MN_RG, MN_SQ, MN_SG, MN_PQ, MN_PG };

template <class M>
struct MnChainOps {
  static constexpr int K = M::K;
  const M& m;
  const MnState& ms;
  const ChainState& st;
  const SamplerParams& sp;
  long long c;
  int lane;
  float v[MN_NV][K];
  float im[K];
  MnChain s;

  __device__ float* vptr(int which, int k) const {
    return ms.vec + (((size_t)c * MN_NV + which) * K + k) * WAVE + lane;
  }
  __device__ void load_all() {
#pragma unroll
    for (int w = 0; w < MN_NV; ++w)
#pragma unroll
      for (int k = 0; k < K; ++k) v[w][k] = *vptr(w, k);
  }
  __device__ void store_all() const {
#pragma unroll
    for (int w = 0; w < MN_NV; ++w)
#pragma unroll
      for (int k = 0; k < K; ++k) *vptr(w, k) = v[w][k];
  }
  __device__ void cp(int dst, int src) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[dst][k] = v[src][k];
  }
  __device__ float* ckp() const { return ms.ck + (size_t)c * (sp.max_depth + 1) * 2 * K * WAVE; }
  __device__ uint32_t itc() const { return sp.run_seed + (uint32_t)s.it; }

  // rows for the batched gradient (api order)
  __device__ void emit(int which) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      if (a >= 0) ms.wq[c * sp.P + a] = v[which][k];
    }
  }
  // gradient and logjoint at the emitted rows; ok as the models' eval defines it
  __device__ float take(bool& ok) {
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      v[MN_G][k] = a >= 0 ? ms.wg[c * sp.P + a] : 0.0f;
      fin = fin && isfinite_f(v[MN_G][k]);
    }
    const float lj = ms.wlj[c];
    ok = wave_all(fin && isfinite_f(lj));
    return lj;
  }

  // iteration start: momentum, H0, the first leaf's half kick and drift
  __device__ void start() {
    const uint32_t ch = (uint32_t)c;
    s.step = st.step[c];
    s.lj0 = s.lj;
    cp(MN_Q0, MN_Q);
    cp(MN_G0, MN_G);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      if (a >= 0) {
        const float z = rand_normal_f(ch, itc(), STREAM_MOMENTUM, (uint32_t)a);
        v[MN_P][k] = z / sqrt_f(im[k]);
      } else {
        v[MN_P][k] = 0.0f;
      }
    }
    s.H0 = kinetic(v[MN_P], im) - (double)s.lj;
    cp(MN_LQ, MN_Q); cp(MN_LP, MN_P); cp(MN_LG, MN_G);
    cp(MN_RQ, MN_Q); cp(MN_RP, MN_P); cp(MN_RG, MN_G);
    cp(MN_PQ, MN_Q); cp(MN_PG, MN_G);
    cp(MN_SQ, MN_Q); cp(MN_SG, MN_G);
    s.ljL = s.lj; s.ljR = s.lj; s.ljP = s.lj; s.ljS = s.lj;
    s.cont_w = -s.H0;
    s.accept_sum = 0.0;
    s.accept_count = 0;
    s.depth = 1;
    s.steps = 0;
    s.divergent = 0;
    s.turning = 0;
    s.prop_energy = s.H0;
    s.prop_error = 0.0;
    s.leaf_counter = 0;
    const float u = readlane(rand_uniform_f(ch, itc(), STREAM_DIRECTION, (uint32_t)lane), 0);
    s.forward = !(u < 0.5f);
    drift();
    s.phase = MN_FIRST;
  }
  __device__ void drift() {
    const float h = (float)s.step;
    const float ss = s.forward ? h : -h;
    const float half = ss / 2.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      v[MN_P][k] = v[MN_P][k] + half * v[MN_G][k];
      v[MN_Q][k] = v[MN_Q][k] + ss * (im[k] * v[MN_P][k]);
    }
    emit(MN_Q);
  }
  __device__ void kick() {
    const float h = (float)s.step;
    const float half = (s.forward ? h : -h) / 2.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) v[MN_P][k] = v[MN_P][k] + half * v[MN_G][k];
  }
  // a doubling round: direction, frontier, subtree state, first leaf
  __device__ void begin_round() {
    const uint32_t ch = (uint32_t)c;
    const float ud = readlane(rand_uniform_f(ch, itc(), STREAM_DIRECTION, (uint32_t)lane), s.depth);
    s.forward = !(ud < 0.5f);
    if (s.forward) { cp(MN_Q, MN_RQ); cp(MN_P, MN_RP); cp(MN_G, MN_RG); }
    else { cp(MN_Q, MN_LQ); cp(MN_P, MN_LP); cp(MN_G, MN_LG); }
    s.cur_lj = s.forward ? s.ljR : s.ljL;
    s.sub_w = -inf_d();
    s.sub_int = 0;
    s.sub_asum = 0.0;
    s.sub_acount = 0;
    s.sub_div = 0;
    s.sub_turn = 0;
    s.sub_energy = 0.0;
    s.sub_error = 0.0;
    s.leaf = 0;
    drift();
    s.phase = MN_LEAF;
  }
  __device__ bool tree_continues() const { return s.depth < sp.max_depth && !s.divergent && !s.turning; }

  // the first leaf's gradient has arrived
  __device__ void post_first() {
    bool ok;
    const float lj1 = take(ok);
    if (!ok) {
      s.divergent = 1;
    } else {
      kick();
      if (s.forward) { cp(MN_RQ, MN_Q); cp(MN_RP, MN_P); cp(MN_RG, MN_G); s.ljR = lj1; }
      else { cp(MN_LQ, MN_Q); cp(MN_LP, MN_P); cp(MN_LG, MN_G); s.ljL = lj1; }
      const double pe = kinetic(v[MN_P], im) - (double)lj1;
      const double delta = pe - s.H0;
      s.steps = 1;
      if (!isfinite_d(delta) || delta > sp.div_thr) {
        s.divergent = 1;
      } else {
        const double cand = -pe;
        double hi, arg;
        lae_split(s.cont_w, cand, hi, arg);
        const double ex = exp_d(lane == 0 ? fmin(0.0, -delta) : arg);
        const double e = readlane(ex, 0);
        s.accept_sum += e < 1.0 ? e : 1.0;
        s.accept_count += 1;
        const double comb = lae_join(hi, readlane(ex, 1));
        const double leaf_lu = log_d(rand_uniform_d((uint32_t)c, itc(), STREAM_LEAF, (uint32_t)lane));
        if (!isfinite_d(s.cont_w) || readlane(leaf_lu, 0) < cand - comb) {
          cp(MN_PQ, MN_Q); cp(MN_PG, MN_G); s.ljP = lj1;
          s.prop_energy = pe;
          s.prop_error = delta;
        }
        s.cont_w = comb;
        s.turning = is_turning_lr(v[MN_LQ], v[MN_LP], v[MN_RQ], v[MN_RP], im) ? 1 : 0;
      }
      s.leaf_counter += 1;
    }
    if (tree_continues()) begin_round();
    else finish();
  }

  // a subtree leaf's gradient has arrived; returns true when the subtree ended
  __device__ bool post_leaf() {
    bool ok;
    const float ljc = take(ok);
    if (!ok) {
      s.sub_div = 1;
      return true;
    }
    kick();
    s.cur_lj = ljc;
    s.sub_int += 1;
    const double pe = kinetic(v[MN_P], im) - (double)ljc;
    const double delta = pe - s.H0;
    if (!isfinite_d(delta) || delta > sp.div_thr) {
      s.sub_div = 1;
      return true;
    }
    const double cand = -pe;
    double hi, arg;
    lae_split(s.sub_w, cand, hi, arg);
    const double ex = exp_d(lane == 0 ? fmin(0.0, -delta) : arg);
    const double e = readlane(ex, 0);
    s.sub_asum += e < 1.0 ? e : 1.0;
    s.sub_acount += 1;
    const double comb = lae_join(hi, readlane(ex, 1));
    const int blk = (int)(s.leaf_counter >> 6);
    const double leaf_lu =
        log_d(rand_uniform_d((uint32_t)c, itc(), STREAM_LEAF, ((uint32_t)blk << 6) + (uint32_t)lane));
    if (!isfinite_d(s.sub_w) || readlane(leaf_lu, (int)(s.leaf_counter & 63u)) < cand - comb) {
      cp(MN_SQ, MN_Q); cp(MN_SG, MN_G); s.ljS = ljc;
      s.sub_energy = pe;
      s.sub_error = delta;
    }
    s.sub_w = comb;
    s.leaf_counter += 1;
    const int li = s.sub_int - 1;
    float* ck = ckp();
    if ((li & 1) == 0) {
      const int slot = __builtin_popcount((unsigned)li);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ck[((slot * 2 + 0) * K + k) * WAVE + lane] = v[MN_Q][k];
        ck[((slot * 2 + 1) * K + k) * WAVE + lane] = v[MN_P][k];
      }
    } else if (dyadic_turning(ck, v[MN_Q], v[MN_P], li, s.forward != 0, im)) {
      s.sub_turn = 1;
      return true;
    }
    s.leaf += 1;
    return s.leaf >= (1 << s.depth);
  }

  // the subtree ended (all leaves, a U-turn or a divergence)
  __device__ void end_round() {
    s.depth += 1;
    if (s.sub_int == 0) {
      s.divergent = s.sub_div;
      finish();
      return;
    }
    if (s.sub_turn || s.sub_div) {
      s.steps += s.sub_int;
      s.accept_sum += s.sub_asum;
      s.accept_count += s.sub_acount;
      s.turning = s.sub_turn;
      s.divergent = s.sub_div;
      finish();
      return;
    }
    if (s.forward) { cp(MN_RQ, MN_Q); cp(MN_RP, MN_P); cp(MN_RG, MN_G); s.ljR = s.cur_lj; }
    else { cp(MN_LQ, MN_Q); cp(MN_LP, MN_P); cp(MN_LG, MN_G); s.ljL = s.cur_lj; }
    double comb_cw = s.cont_w;
    if (isfinite_d(s.sub_w)) {
      comb_cw = logaddexp_d(s.cont_w, s.sub_w);
      const double merge_lu = log_d(rand_uniform_d((uint32_t)c, itc(), STREAM_MERGE, (uint32_t)lane));
      if (readlane(merge_lu, s.depth - 1) < s.sub_w - s.cont_w) {
        cp(MN_PQ, MN_SQ); cp(MN_PG, MN_SG); s.ljP = s.ljS;
        s.prop_energy = s.sub_energy;
        s.prop_error = s.sub_error;
      }
    }
    s.steps += s.sub_int;
    s.accept_sum += s.sub_asum;
    s.accept_count += s.sub_acount;
    if (isfinite_d(s.sub_w)) s.cont_w = comb_cw;
    s.turning = (s.sub_turn || is_turning_lr(v[MN_LQ], v[MN_LP], v[MN_RQ], v[MN_RP], im)) ? 1 : 0;
    if (tree_continues()) begin_round();
    else finish();
  }

  // the iteration's transition is complete: accept, adapt, record, next iteration
  __device__ void finish() {
    const bool moved = any_differs(v[MN_PQ], v[MN_Q0]);
    const double accept_prob = s.accept_count == 0 ? 0.0 : s.accept_sum / (double)s.accept_count;
    if (moved) {
      cp(MN_Q, MN_PQ); cp(MN_G, MN_PG); s.lj = s.ljP;
    } else {
      cp(MN_Q, MN_Q0); cp(MN_G, MN_G0); s.lj = s.lj0;
    }
    const uint8_t flags = (uint8_t)((s.divergent ? FLAG_DIVERGENT : 0) | (moved ? FLAG_ACCEPTED : 0));
    const int it = s.it;
    const int P = sp.P;
    if (it <= sp.num_warmup) {
      if (sp.adapt_step) {
        double dual[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) dual[i] = st.dual[c * 5 + i];
        double step = dual_update(dual, s.divergent ? 0.0 : accept_prob, sp.target_accept);
        if (it == sp.num_warmup) step = exp_d(dual[2]);
        if (lane == 0) {
          st.step[c] = step;
#pragma unroll
          for (int i = 0; i < 5; ++i) st.dual[c * 5 + i] = dual[i];
        }
      }
      const long long row = (long long)(it - 1) * sp.C + c;
      store_vec(m, sp.hist_pos + row * P, v[MN_Q]);
      if (lane == 0) {
        sp.hist_acc[row] = accept_prob;
        sp.hist_flags[row] = flags;
      }
    } else {
      const long long srow = c * sp.num_samples + (it - sp.num_warmup - 1);
      store_vec(m, sp.draws + srow * P, v[MN_Q]);
      if (lane == 0) {
        sp.o_lj[srow] = (double)s.lj;
        sp.o_acc[srow] = accept_prob;
        sp.o_energy[srow] = s.prop_energy;
        sp.o_err[srow] = s.prop_error;
        sp.o_accepted[srow] = moved ? 1 : 0;
        sp.o_divergent[srow] = s.divergent ? 1 : 0;
        sp.o_depth[srow] = s.depth;
        sp.o_steps[srow] = s.steps;
        atomicAdd(sp.eval_counter, (unsigned long long)s.steps);
      }
    }
    store_vec(m, st.pos + c * P, v[MN_Q]);
    store_vec(m, st.grad + c * P, v[MN_G]);
    if (lane == 0) st.lj[c] = s.lj;
    s.it += 1;
    s.phase = MN_DONE;
  }
};

template <class M>
__global__ __launch_bounds__(64) void k_mn_step(M model, ChainState st, SamplerParams sp, MnState ms, int it_end) {
  MnChainOps<M> o{model, ms, st, sp, (long long)blockIdx.x, (int)threadIdx.x};
  o.s = ms.ch[o.c];
  if (o.s.phase == MN_DONE) return;
  constexpr int K = M::K;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = model.api(o.lane, k);
    o.im[k] = a >= 0 ? st.im_f[a] : 0.0f;
  }
  if (o.s.phase == MN_START) {
    load_chain(model, st, sp.P, o.c, o.v[MN_Q], o.v[MN_G], o.im);
    o.s.lj = st.lj[o.c];
    o.start();
  } else {
    o.load_all();
    if (o.s.phase == MN_FIRST) o.post_first();
    else if (o.post_leaf()) o.end_round();
    else o.drift();
    if (o.s.phase == MN_DONE && o.s.it <= it_end) o.start();
  }
  o.store_all();
  if (o.lane == 0) {
    ms.ch[o.c] = o.s;
    if (o.s.phase != MN_DONE) atomicAdd(ms.active, 1);
  }
}

__global__ void k_mn_reset(MnState ms, int C, int it_begin) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  ms.ch[c].phase = MN_START;
  ms.ch[c].it = it_begin;
}

// ---- per-chain step search in lock step (k_step_search restated): one trial
// leapfrog step per call, the gradient from the batched engine
enum { MS_IDLE = 0, MS_TRIAL = 1 };

template <class M>
struct MsOps {
  static constexpr int K = M::K;
  const M& m;
  const MnState& ms;
  const ChainState& st;
  long long c;
  int lane, P;
  float q[K], g[K], p[K], im[K];
  MnChain s;

  __device__ float* pptr(int k) const { return ms.vec + (((size_t)c * MN_NV + MN_P) * K + k) * WAVE + lane; }
  // momentum, H0, half kick, drift of a trial at step x; rows to ms.wq
  __device__ void trial_pre(double x, uint32_t itc) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      if (a >= 0) {
        const float z = rand_normal_f((uint32_t)c, itc, STREAM_STEP_SEARCH, (uint32_t)a);
        p[k] = z / sqrt_f(im[k]);
      } else {
        p[k] = 0.0f;
      }
    }
    s.search_H0 = kinetic(p, im) - (double)st.lj[c];
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) fin = fin && isfinite_f(g[k]);
    const bool valid = wave_all(fin);
    const float h = (float)x;
    const float half = 0.5f * h;
    float nq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (valid) p[k] = p[k] + half * g[k];
      nq[k] = q[k] + (h * im[k]) * p[k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      if (a >= 0) ms.wq[c * P + a] = nq[k];
      *pptr(k) = p[k];
    }
    s.search_valid = valid ? 1 : 0;
    s.search_next = x;
    s.phase = MS_TRIAL;
  }
  __device__ double trial_post() {
    float ng[K];
    bool fin = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      ng[k] = a >= 0 ? ms.wg[c * P + a] : 0.0f;
      fin = fin && isfinite_f(ng[k]);
      p[k] = *pptr(k);
    }
    const float nlj = ms.wlj[c];
    const bool ok = wave_all(fin && isfinite_f(nlj));
    const bool valid = s.search_valid && ok;
    if (!valid) return -inf_d();
    const float half = 0.5f * (float)s.search_next;
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = -(p[k] + half * ng[k]);
    const double H1 = kinetic(p, im) - (double)nlj;
    return s.search_H0 - H1;
  }
  __device__ void done() {
    double x = s.search_s;
    x = x < 1e-8 ? 1e-8 : (x > 1e3 ? 1e3 : x);
    if (lane == 0) {
      st.step[c] = x;
      const double ls = log_d(x);
      st.dual[c * 5 + 0] = log_d(10.0 * x);
      st.dual[c * 5 + 1] = ls;
      st.dual[c * 5 + 2] = ls;
      st.dual[c * 5 + 3] = 0.0;
      st.dual[c * 5 + 4] = 0.0;
    }
    s.phase = MS_IDLE;
  }
  __device__ void load() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int a = m.api(lane, k);
      q[k] = a >= 0 ? st.pos[c * P + a] : 0.0f;
      g[k] = a >= 0 ? st.grad[c * P + a] : 0.0f;
      im[k] = a >= 0 ? st.im_f[a] : 0.0f;
    }
  }
};

template <class M>
__global__ __launch_bounds__(64) void k_mn_search_begin(M model, ChainState st, MnState ms, int P, uint32_t itc,
                                                        int do_search) {
  MsOps<M> o{model, ms, st, (long long)blockIdx.x, (int)threadIdx.x, P};
  o.s = ms.ch[o.c];
  o.s.search_s = st.step[o.c];
  if (!do_search) {
    o.done();
  } else {
    o.load();
    o.s.search_t = -1;  // the first trial at the current step
    o.trial_pre(o.s.search_s, itc);
  }
  if (o.lane == 0) {
    ms.ch[o.c] = o.s;
    if (o.s.phase != MS_IDLE) atomicAdd(ms.active, 1);
  }
}

template <class M>
__global__ __launch_bounds__(64) void k_mn_search_step(M model, ChainState st, MnState ms, int P, uint32_t itc) {
  MsOps<M> o{model, ms, st, (long long)blockIdx.x, (int)threadIdx.x, P};
  o.s = ms.ch[o.c];
  if (o.s.phase == MS_IDLE) return;
  o.load();
  const double log_target = -0.6931471805599453;  // log(0.5)
  const double lar = o.trial_post();
  bool stop = false;
  if (o.s.search_t < 0) {
    o.s.search_dir = lar > log_target ? 1.0 : -1.0;
    o.s.search_t = 0;
  } else {
    o.s.search_s = o.s.search_next;
    const double dir = o.s.search_dir;
    if ((dir > 0.0 && lar <= log_target) || (dir < 0.0 && lar >= log_target)) stop = true;
    o.s.search_t += 1;
    if (o.s.search_t >= 20) stop = true;
  }
  if (!stop) {
    const double next = o.s.search_s * (o.s.search_dir > 0.0 ? 2.0 : 0.5);
    if (next < 1e-8 || next > 1e3) stop = true;
    else o.trial_pre(next, itc);
  }
  if (stop) o.done();
  if (o.lane == 0) {
    ms.ch[o.c] = o.s;
    if (o.s.phase != MS_IDLE) atomicAdd(ms.active, 1);
  }
}

// wavefront models keep their own vector layout; the batched-only plans the
// identity layout of K slots
template <class F>
static hipError_t dispatch_layout(const DevModelDesc& d, F&& f) {
  if (d.kind == MODEL_GLM_MFMA || d.kind == MODEL_GP) {
    const int K = masked_slots(d);
    IdentityLayout<16> l16{d.P};
    if (K == 1) return f(IdentityLayout<1>{d.P});
    if (K == 2) return f(IdentityLayout<2>{d.P});
    if (K == 4) return f(IdentityLayout<4>{d.P});
    if (K == 8) return f(IdentityLayout<8>{d.P});
    if (K == 16) return f(l16);
    return hipErrorInvalidValue;
  }
  return dispatch(d, std::forward<F>(f));
}

int masked_slots(const DevModelDesc& d) {
  if (d.kind != MODEL_GLM_MFMA && d.kind != MODEL_GP) return d.K;
  const int k = (d.P + WAVE - 1) / WAVE;
  for (int s = 1; s <= 16; s *= 2)
    if (k <= s) return s;
  return -1;
}

hipError_t launch_mn_reset(const MnState& ms, int C, int it_begin, hipStream_t stream) {
  hipLaunchKernelGGL(k_mn_reset, dim3((C + 255) / 256), dim3(256), 0, stream, ms, C, it_begin);
  return hipGetLastError();
}

hipError_t launch_mn_step(const DevModelDesc& d, const ChainState& st, const SamplerParams& sp, const MnState& ms,
                          int it_end, hipStream_t stream) {
  return dispatch_layout(d, [&](auto m) {
    hipLaunchKernelGGL((k_mn_step<decltype(m)>), dim3(sp.C), dim3(WAVE), 0, stream, m, st, sp, ms, it_end);
    return hipGetLastError();
  });
}

hipError_t launch_mn_search_begin(const DevModelDesc& d, const ChainState& st, const MnState& ms, int C, int P,
                                  uint32_t itc, int do_search, hipStream_t stream) {
  return dispatch_layout(d, [&](auto m) {
    hipLaunchKernelGGL((k_mn_search_begin<decltype(m)>), dim3(C), dim3(WAVE), 0, stream, m, st, ms, P, itc,
                       do_search);
    return hipGetLastError();
  });
}

hipError_t launch_mn_search_step(const DevModelDesc& d, const ChainState& st, const MnState& ms, int C, int P,
                                 uint32_t itc, hipStream_t stream) {
  return dispatch_layout(d, [&](auto m) {
    hipLaunchKernelGGL((k_mn_search_step<decltype(m)>), dim3(C), dim3(WAVE), 0, stream, m, st, ms, P, itc);
    return hipGetLastError();
  });
}

hipError_t launch_mass_replay(const float* hist_pos, const double* hist_acc,
                              const uint8_t* hist_flags, int C, int P, int it_lo, int it_hi,
                              int wl, int use_accepted, double reg, int min_samples,
                              double* im_d, float* im_f, double* summary, hipStream_t stream) {
  if (P > REPLAY_THREADS * REPLAY_PER_THREAD) return hipErrorInvalidValue;
  // staged block within 64 KB of LDS: per chain 5 scalars of 8 bytes and P positions
  const int cb = std::max(1, std::min(256, (65536 - 64) / (4 * P + 40)));
  const size_t shmem = (size_t)cb * 40 + 32 + (size_t)cb * P * sizeof(float);
  if (P <= REPLAY_THREADS)
    hipLaunchKernelGGL(k_mass_replay<1>, dim3(1), dim3(REPLAY_THREADS), shmem, stream, hist_pos, hist_acc,
                       hist_flags, C, P, it_lo, it_hi, wl, use_accepted, reg, min_samples, im_d,
                       im_f, summary, cb);
  else
    hipLaunchKernelGGL(k_mass_replay<REPLAY_PER_THREAD>, dim3(1), dim3(REPLAY_THREADS), shmem, stream, hist_pos,
                       hist_acc, hist_flags, C, P, it_lo, it_hi, wl, use_accepted, reg, min_samples, im_d,
                       im_f, summary, cb);
  return hipGetLastError();
}

hipError_t launch_philox_fill(uint32_t chain, uint32_t iteration, uint32_t stream_id,
                              uint32_t first, int64_t n, uint32_t* blocks, float* uf, float* nf,
                              double* ud, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int threads = 256;
  const long long grid = (n + threads - 1) / threads;
  hipLaunchKernelGGL(k_philox_fill, dim3((unsigned)grid), dim3(threads), 0, stream, chain,
                     iteration, stream_id, first, (long long)n, blocks, uf, nf, ud);
  return hipGetLastError();
}

}  // namespace ut
