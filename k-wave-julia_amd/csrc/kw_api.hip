// C ABI implementation: host-side FP64 table generation, device buffers,
// launch sequencing of one PSTD time step, sensor read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/kwave_hip.h"
#include "kw_host.hpp"
#include "kw_launch.hpp"

using namespace kw;
using namespace kwh;

namespace {

// ---------------------------------------------------------------------------
// Launch-shape tuning table: lines per workgroup and plan family for each
// (precision bytes, pass class, line length), filled by kw_autotune's timed
// trials or by kw_tune_import; read when a simulation or batch is created.
// ---------------------------------------------------------------------------
enum TuneClass { TC_DIR_X = 0, TC_DIR_Y, TC_DIR_Z, TC_SPEC_Y, TC_SPEC_Z, TC_C2R, TC_COUNT };
std::mutex g_tune_mu;
std::map<std::array<int, 3>, int> g_tune;      // lines per workgroup
std::map<std::array<int, 3>, int> g_tune_pv;   // plan family (absent: the default)
// Inside an autotuner trial: lines per workgroup = default >> g_tune_shift
// (shift -1 doubles it, up to the pass's maximum) and plan family
// g_tune_plan where the line length has it (-1: defaults); kNoTrial outside
// trials.
constexpr int kNoTrial = 99;
thread_local int g_tune_shift = kNoTrial;
thread_local int g_tune_plan = kNoTrial;

// Plan family of a single simulation's pass: the trial's or the table's
// choice where the line length has alternatives, else -1 (the default).
int tuned_plan(int prec, int cls, int n) {
  if (g_tune_shift != kNoTrial) return plan_alternative(n, g_tune_plan) ? g_tune_plan : -1;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune_pv.find({prec, cls, n});
  return it != g_tune_pv.end() && plan_alternative(n, it->second) ? it->second : -1;
}

// Lines per workgroup of a pass: the tuning table's entry for (precision,
// class, line length) or the trial's choice, clamped to [1, maxl] (maxl: the
// most lines that fit the pass's thread limit and the workgroup LDS).
int tuned_lpw(int prec, int cls, int n, int dflt, int maxl) {
  if (g_tune_shift != kNoTrial) {
    const int v = g_tune_shift >= 0 ? dflt >> g_tune_shift : dflt << (-g_tune_shift);
    return std::max(1, std::min(maxl, v));
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find({prec, cls, n});
  return it == g_tune.end() ? dflt : std::min(maxl, std::max(1, it->second));
}

// ---------------------------------------------------------------------------
// Simulation
// ---------------------------------------------------------------------------
struct SimBase {
  virtual ~SimBase() = default;
  virtual void run(long long n) = 0;
  virtual void sync() = 0;
  virtual long long steps_done() const = 0;
  virtual void get_field(int which, double* out) = 0;
  virtual void set_field(int which, const double* in) = 0;
  virtual void sensor_shape(uint32_t f, long long* r, long long* c) = 0;
  virtual void get_sensor(uint32_t f, double* out) = 0;
  virtual void profile(long long n, int maxc, int* nc, const char** names, double* ms, long long* launches) = 0;
  virtual void grad_c0(const double* obs, long long ckpt, double* grad, double* loss) = 0;
};

struct SrcDev {
  DevBuf off, pos, row;
  bool any = false;
};

template <typename T>
class Sim final : public SimBase {
 public:
  // batch_mode: fixed workgroup shapes per line length so that a batch can
  // launch the passes of many simulations together (see Batch below).
  // nfire > 1 (2-D only): `d` holds nfire firings of one grid and medium
  // stacked along a third, uncoupled axis (maps replicated per firing,
  // source/sensor indices offset by the firing's plane); every pass runs over
  // all firings at once and the EOS transform stays 2-D per firing.
  Sim(int device, const kw_sim_desc& d, bool batch_mode = false, int nfire = 1)
      : dev_(device), batch_mode_(batch_mode), nfire_(nfire) {
    build(d);
  }
  ~Sim() override {
    if (stream_) {
      (void)hipSetDevice(dev_);
      (void)hipStreamSynchronize(stream_);
      (void)hipStreamDestroy(stream_);
    }
  }

  void run(long long n) override {
    HIPC(hipSetDevice(dev_));
    for (long long i = 0; i < n && t_ < Nt_; ++i) step(nullptr);
    HIPC(hipGetLastError());
    if (t_ == Nt_ && !finalized_) finalize();
  }
  void sync() override {
    HIPC(hipSetDevice(dev_));
    HIPC(hipStreamSynchronize(stream_));
  }
  long long steps_done() const override { return t_; }

  // device buffer of a field selector (kw_field), or null if the selector is invalid here
  DevBuf* field_buf(int which) {
    if (which == KW_FIELD_P) return &p_;
    if (which >= KW_FIELD_UX && which <= KW_FIELD_UZ && which - KW_FIELD_UX < nd_) return &u_[which - KW_FIELD_UX];
    if (which >= KW_FIELD_RHOX && which <= KW_FIELD_RHOZ && which - KW_FIELD_RHOX < nd_)
      return &rho_[which - KW_FIELD_RHOX];
    if (which == KW_FIELD_RHO_PREV && stokes_ && rho_prev_.p) return &rho_prev_;
    return nullptr;
  }

  void get_field(int which, double* out) override {
    HIPC(hipSetDevice(dev_));
    DevBuf* b = field_buf(which);
    if (!b) raise(KW_ERR_INVALID, "invalid field selector");
    copy_out(b->as<T>(), nvox_, out);
  }

  void set_field(int which, const double* in) override {
    HIPC(hipSetDevice(dev_));
    DevBuf* b = field_buf(which);
    if (!b) raise(KW_ERR_INVALID, "invalid field selector");
    HIPC(hipStreamSynchronize(stream_));
    std::vector<T> h(nvox_);
    for (long long i = 0; i < nvox_; ++i) h[i] = (T)in[i];
    HIPC(hipMemcpy(b->p, h.data(), nvox_ * sizeof(T), hipMemcpyHostToDevice));
  }

  void sensor_shape(uint32_t f, long long* r, long long* c) override {
    if (!(rec_ & f) && !(tr_ && f == KW_REC_P_FINAL)) raise(KW_ERR_INVALID, "field not recorded");
    *c = 1;
    *r = nsens_;
    if (f == KW_REC_P || f == KW_REC_UX || f == KW_REC_UY || f == KW_REC_UZ) *c = Nt_alloc_;
    if (tr_ && f == KW_REC_P_FINAL) *r = nvox_;
  }

  void get_sensor(uint32_t f, double* out) override {
    HIPC(hipSetDevice(dev_));
    sync();
    long long rows, cols;
    sensor_shape(f, &rows, &cols);
    if (tr_ && f == KW_REC_P_FINAL) {
      copy_out(p_.as<T>(), nvox_, out);
      return;
    }
    if (cartesian_) {
      // Cartesian masks are preallocated and never filled (first_order.jl:417-421)
      double init = 0.0;
      if (f == KW_REC_P_MAX || f == KW_REC_U_MAX || f == KW_REC_I_MAX) init = -INFINITY;
      if (f == KW_REC_P_MIN) init = INFINITY;
      for (long long i = 0; i < rows * cols; ++i) out[i] = init;
      return;
    }
    switch (f) {
      case KW_REC_P:
        if ((dir_ || fr_) && t_ == Nt_) {
          apply_post_filters();
          HIPC(hipMemcpy(out, pfilt_.p, rows * cols * sizeof(double), hipMemcpyDeviceToHost));
        } else {
          copy_out(p_ts_.as<T>(), rows * cols, out);
        }
        break;
      case KW_REC_UX:
      case KW_REC_UY:
      case KW_REC_UZ: {
        const int d = f == KW_REC_UX ? 0 : (f == KW_REC_UY ? 1 : 2);
        if (d >= nd_) {
          for (long long i = 0; i < rows * cols; ++i) out[i] = 0.0;
        } else {
          copy_out(u_ts_[d].as<T>(), rows * cols, out);
        }
        break;
      }
      case KW_REC_P_MAX:
        copy_out(p_max_.as<T>(), rows, out);
        break;
      case KW_REC_P_MIN:
        copy_out(p_min_.as<T>(), rows, out);
        break;
      case KW_REC_P_FINAL:
        copy_out(p_final_.as<T>(), rows, out);
        break;
      case KW_REC_P_RMS:
        HIPC(hipMemcpy(out, p_rms_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      case KW_REC_U_MAX:
        HIPC(hipMemcpy(out, u_max_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      case KW_REC_U_RMS:
        HIPC(hipMemcpy(out, u_rms_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      case KW_REC_U_FINAL:
        HIPC(hipMemcpy(out, u_final_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      case KW_REC_I_AVG:
        HIPC(hipMemcpy(out, I_avg_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      case KW_REC_I_MAX:
        HIPC(hipMemcpy(out, I_max_.p, rows * sizeof(double), hipMemcpyDeviceToHost));
        break;
      default:
        raise(KW_ERR_INVALID, "unknown sensor field");
    }
  }

  void profile(long long n, int maxc, int* nc, const char** names, double* ms, long long* launches) override {
    HIPC(hipSetDevice(dev_));
    Prof pr;
    pr.ev.reserve(64);
    for (long long i = 0; i < n && t_ < Nt_; ++i) step(&pr);
    HIPC(hipStreamSynchronize(stream_));
    std::vector<double> tot(kNumClasses, 0.0);
    std::vector<long long> cnt(kNumClasses, 0);
    for (auto& r : pr.ev) {
      float e = 0.f;
      HIPC(hipEventElapsedTime(&e, r.a, r.b));
      tot[r.cls] += e;
      cnt[r.cls] += 1;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    int k = 0;
    for (int c = 0; c < kNumClasses && k < maxc; ++c) {
      if (cnt[c] == 0) continue;
      names[k] = kClassNames[c];
      ms[k] = tot[c] / (double)cnt[c];
      launches[k] = cnt[c];
      ++k;
    }
    *nc = k;
    if (t_ == Nt_ && !finalized_) finalize();
  }

  // Autotuner trial: `warm` untimed steps, then per-class device time summed
  // over `steps` timed steps (ms per step), with the lines per workgroup and
  // the line length each class ran with (n = 0: class absent).
  void trial(long long warm, long long steps, double* ms, int* lp, int* nn, int* pv) {
    HIPC(hipSetDevice(dev_));
    for (long long i = 0; i < warm && t_ < Nt_; ++i) step(nullptr);
    Prof pr;
    long long done = 0;
    for (; done < steps && t_ < Nt_; ++done) step(&pr);
    HIPC(hipStreamSynchronize(stream_));
    static const int cls_of[kNumClasses] = {TC_DIR_X, TC_DIR_Y,  TC_DIR_Z, TC_SPEC_Y, TC_SPEC_Z, TC_SPEC_Y,
                                            TC_SPEC_Y, TC_C2R,    -1,       -1,        -1};
    static_assert(kNumClasses == 11, "one tuning class (or -1) per pass class");
    for (int c = 0; c < TC_COUNT; ++c) ms[c] = 0.0;
    for (auto& r : pr.ev) {
      float e = 0.f;
      HIPC(hipEventElapsedTime(&e, r.a, r.b));
      if (cls_of[r.cls] >= 0) ms[cls_of[r.cls]] += e;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (int c = 0; c < TC_COUNT; ++c) ms[c] /= (double)std::max<long long>(done, 1);
    const PassCfg* cf[TC_COUNT] = {&cfg_[0], nd_ > 1 ? &cfg_[1] : nullptr, nd_ > 2 ? &cfg_[2] : nullptr,
                                   nd_ > 1 ? &spec_[0] : nullptr, nd_ > 2 ? &spec_[1] : nullptr, &c2r_};
    for (int c = 0; c < TC_COUNT; ++c) {
      lp[c] = cf[c] ? cf[c]->req : 0;
      nn[c] = cf[c] ? cf[c]->n : 0;
      pv[c] = cf[c] ? cf[c]->pv : -1;
    }
  }

  // Loss 1/2 sum (p - obs)^2 over the recorded sensor p and its gradient
  // w.r.t. the c0 map by the discrete adjoint with checkpoints every `ckpt`
  // steps (forward states saved on the way out, each segment recomputed on
  // the way back with sum_d rho_d stored per step).  obs is n_sensor x Nt,
  // column-major; grad is the full grid (column-major, Float64).
  void grad_c0(const double* obs, long long ckpt, double* grad, double* loss) override {
    HIPC(hipSetDevice(dev_));
    if (!het_c_) raise(KW_ERR_UNSUPPORTED, "the c0 gradient needs a sound-speed map");
    if (nonlinear_) raise(KW_ERR_UNSUPPORTED, "the c0 gradient supports the linear equation of state");
    if (axisym_) raise(KW_ERR_UNSUPPORTED, "the c0 gradient is not available for the axisymmetric solver");
    if (stokes_) raise(KW_ERR_UNSUPPORTED, "the c0 gradient needs the heterogeneous equation-of-state branch");
    if (maps_f16_) raise(KW_ERR_UNSUPPORTED, "the c0 gradient needs maps in the kernel precision");
    for (int a = 0; a < nd_; ++a)
      if (bs_[a]) raise(KW_ERR_UNSUPPORTED, "the c0 gradient needs axis lengths with a compiled line-FFT plan");
    if (tr_ || (psrc_[0].any && p_dir_)) raise(KW_ERR_UNSUPPORTED, "time reversal / Dirichlet sources");
    for (int a = 0; a < nd_; ++a)
      if (usrc_[a].any && u_dir_) raise(KW_ERR_UNSUPPORTED, "Dirichlet velocity sources");
    if (!(rec_ & KW_REC_P) || cartesian_ || nsens_ == 0) raise(KW_ERR_INVALID, "the loss needs recorded p at a binary mask");
    if (t_ != 0) raise(KW_ERR_STATE, "grad_c0 must start from step 0");
    if (ckpt < 1) ckpt = 1;
    const long long Nt = Nt_, nseg = (Nt + ckpt - 1) / ckpt;
    const size_t fb = nvox_ * sizeof(T);
    const int nf = 1 + 2 * nd_;
    // ---- forward with checkpoints
    DevBuf ck;
    ck.alloc((size_t)nseg * nf * fb);
    auto fields = [&](int f) -> T* {
      if (f == 0) return p_.as<T>();
      if (f <= nd_) return u_[f - 1].as<T>();
      return rho_[f - 1 - nd_].as<T>();
    };
    auto save = [&](long long s) {
      for (int f = 0; f < nf; ++f)
        HIPC(hipMemcpyAsync(ck.as<char>() + ((size_t)s * nf + f) * fb, fields(f), fb, hipMemcpyDeviceToDevice, stream_));
    };
    auto restore = [&](long long s) {
      for (int f = 0; f < nf; ++f)
        HIPC(hipMemcpyAsync(fields(f), ck.as<char>() + ((size_t)s * nf + f) * fb, fb, hipMemcpyDeviceToDevice, stream_));
    };
    for (long long s = 0; s < nseg; ++s) {
      save(s);
      for (long long i = 0; i < ckpt && t_ < Nt; ++i) step(nullptr);
    }
    HIPC(hipStreamSynchronize(stream_));
    // ---- residuals and loss
    const long long nres = nsens_ * Nt;
    DevBuf dobs, dres, dloss, dg, dgp, dgref;
    upload(dobs, std::vector<double>(obs, obs + nres));
    dres.alloc(nres * sizeof(double));
    const unsigned nl = adj_reduce_blocks(nres), ng = adj_reduce_blocks(plane_);
    upload(dloss, std::vector<double>(nl, 0.0));
    upload(dg, std::vector<double>(nvox_, 0.0));    // source and final terms, per stacked voxel
    upload(dgp, std::vector<double>(plane_, 0.0));  // per-step term, summed over the firings
    upload(dgref, std::vector<double>(ng, 0.0));
    {
      const void* pp[4] = {p_ts_.p, dobs.p, dres.p, dloss.p};
      HIPC(launch_adj_small<T>(2, stream_, pp, nullptr, nres, nd_));
    }
    // ---- adjoint buffers
    DevBuf Pb, Q, Qd, RS, Ub[3], Rb[3], dsig, didx, dfac;
    Pb.alloc(fb);
    Q.alloc(fb);
    Qd.alloc(fb);
    RS.alloc((size_t)ckpt * fb);
    HIPC(hipMemsetAsync(Pb.p, 0, fb, stream_));
    for (int a = 0; a < nd_; ++a) {
      Ub[a].alloc(fb);
      Rb[a].alloc(fb);
      HIPC(hipMemsetAsync(Ub[a].p, 0, fb, stream_));
      HIPC(hipMemsetAsync(Rb[a].p, 0, fb, stream_));
    }
    if (!psrc_idx_h_.empty()) {
      upload(didx, psrc_idx_h_);
      upload(dfac, psrc_fac_h_);
      upload(dsig, psrc_sig_h_);   // n_src x ncols, column-major
    }
    // dK/dc_ref multiplier table (same octant layout as M1, 1/N folded in)
    DevBuf M2d;
    build_dkappa(M2d);
    // adjoint pass arguments per axis
    AdjArgs<T> aa[3];
    for (int a = 0; a < nd_; ++a) {
      AdjArgs<T>& A = aa[a];
      std::memset(&A, 0, sizeof(A));
      A.Rb = Rb[a].as<T>();
      A.Ub = Ub[a].as<T>();
      A.Pb = Pb.as<T>();
      A.Q = Q.as<T>();
      A.c0sq = c0sq_.as<T>();
      A.rho0 = pa_[a].rho0;
      A.rho0_s = pa_[a].rho0_s;
      A.pml = pa_[a].pml;
      A.pml_sg = pa_[a].pml_sg;
      A.Dp = pa_[a].Dp;
      A.Dm = pa_[a].Dm;
      A.tw = pa_[a].tw;
      A.dt = pa_[a].dt;
      A.Nx = pa_[a].Nx;
      A.Sa = pa_[a].Sa;
      A.So = pa_[a].So;
      A.n_lines = pa_[a].n_lines;
      A.x_blocks = pa_[a].x_blocks;
    }
    // ---- reverse sweep over segments
    for (long long s = nseg - 1; s >= 0; --s) {
      const long long t0 = s * ckpt, t1 = std::min(Nt, t0 + ckpt);
      restore(s);
      t_ = t0;
      no_record_ = true;
      for (long long n = t0 + 1; n <= t1; ++n) {
        pa_[0].rho_sum_out = RS.as<T>() + (size_t)(n - t0 - 1) * nvox_;
        step(nullptr);
      }
      pa_[0].rho_sum_out = nullptr;
      no_record_ = false;
      for (long long n = t1; n > t0; --n) {
        const T* rs = RS.as<T>() + (size_t)(n - t0 - 1) * nvox_;
        // Pbar += M^T r_n
        {
          const void* pp[3] = {Pb.p, sidx_.p, dres.as<double>() + (size_t)(n - 1) * nsens_};
          HIPC(launch_adj_small<T>(3, stream_, pp, nullptr, nsens_, nd_));
        }
        apply_K(Pb.as<T>(), M2d.as<T>(), Q.as<T>(), Qd.as<T>());
        {
          const void* pp[6] = {c0sq_.p, rs, Q.p, Qd.p, dgp.p, dgref.p};
          HIPC(launch_adj_small<T>(0, stream_, pp, nullptr, plane_, nd_, nfire_));
        }
        if (!psrc_idx_h_.empty()) {
          const long long ns = (long long)psrc_idx_h_.size();
          const long long col = std::min<long long>(n, p_nc_) - 1;
          const void* pp[9] = {Rb[0].p, nd_ > 1 ? Rb[1].p : nullptr, nd_ > 2 ? Rb[2].p : nullptr, c0sq_.p, Q.p,
                               didx.p, dfac.p, dsig.as<double>() + col * ns, dg.p};
          HIPC(launch_adj_small<T>(1, stream_, pp, nullptr, ns, nd_));
        }
        for (int a = 0; a < nd_; ++a) {
          aa[a].accumulate = a > 0 ? 1 : 0;
          HIPC(launch_adj_pass<T>(cfg_[a].n, a == 0, cfg_[a].grid, cfg_[a].block, cfg_[a].shmem, stream_, aa[a],
                                  cfg_[a].pv));
        }
      }
    }
    {
      // block partial sums added in a fixed order: results do not depend on scheduling
      std::vector<double> part(ng);
      HIPC(hipStreamSynchronize(stream_));
      HIPC(hipMemcpy(part.data(), dgref.p, ng * sizeof(double), hipMemcpyDeviceToHost));
      double gsum = 0.0;
      for (double v : part) gsum += v;
      // c_ref = mean(c0) over one firing's grid; with stacked firings each copy
      // of a voxel takes 1/nfire of the term, so the sum over firings adds it once
      const double gterm = gsum / (double)plane_ / (double)nfire_;
      const void* pp[6] = {dg.p, has_p0_ ? p0_keep_.p : nullptr, c0sq_.p, Rb[0].p,
                           nd_ > 1 ? Rb[1].p : nullptr, nd_ > 2 ? Rb[2].p : nullptr};
      HIPC(launch_adj_small<T>(4, stream_, pp, &gterm, nvox_, nd_));
    }
    HIPC(hipStreamSynchronize(stream_));
    {
      // the per-step term (already summed over the firings) plus each
      // firing's source and final terms, added in firing order
      std::vector<double> all(nvox_), gp(plane_);
      HIPC(hipMemcpy(all.data(), dg.p, nvox_ * sizeof(double), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(gp.data(), dgp.p, plane_ * sizeof(double), hipMemcpyDeviceToHost));
      for (long long i = 0; i < plane_; ++i) grad[i] = gp[i] + all[i];
      for (int f = 1; f < nfire_; ++f)
        for (long long i = 0; i < plane_; ++i) grad[i] += all[(size_t)f * plane_ + i];
    }
    {
      std::vector<double> part(nl);
      HIPC(hipMemcpy(part.data(), dloss.p, nl * sizeof(double), hipMemcpyDeviceToHost));
      double l = 0.0;
      for (double v : part) l += v;
      *loss = l;
    }
    t_ = Nt_;
    if (!finalized_) finalize();
  }

 private:
  // Q = K f and Qd = dK/dc_ref f for a real field f (K = F^-1 M1 F)
  void apply_K(const T* f, const T* M2, T* Q, T* Qd) {
    HIPC(launch_r2c<T>(c2r_.n, c2r_.grid, c2r_.block, c2r_.shmem, stream_, f, S1_.as<Cx<T>>(), (int)(N_[1] * N_[2]),
                       (int)NHP_, tw_[0].as<Cx<T>>(), c2r_.pv));
    if (!S2_.p) S2_.alloc((size_t)NHP_ * N_[1] * N_[2] * sizeof(Cx<T>));
    if (nd_ == 3) {
      SpecArgs<T> Y = sa_[0];
      Y.mode = 0;
      Y.S2 = nullptr;
      HIPC(launch_spec<T>(spec_[0].n, spec_[0].grid, spec_[0].block, spec_[0].shmem, stream_, Y, spec_[0].pv));
      SpecArgs<T> Z = sa_[1];
      Z.mode = 5;
      Z.S2 = S2_.as<Cx<T>>();
      Z.M2 = M2;
      HIPC(launch_spec<T>(spec_[1].n, spec_[1].grid, spec_[1].block, spec_[1].shmem, stream_, Z, spec_[1].pv));
      Y.mode = 1;
      HIPC(launch_spec<T>(spec_[0].n, spec_[0].grid, spec_[0].block, spec_[0].shmem, stream_, Y, spec_[0].pv));
      Y.S1 = S2_.as<Cx<T>>();
      HIPC(launch_spec<T>(spec_[0].n, spec_[0].grid, spec_[0].block, spec_[0].shmem, stream_, Y, spec_[0].pv));
    } else if (nd_ == 2) {
      SpecArgs<T> Y = sa_[0];
      Y.mode = 5;
      Y.S2 = S2_.as<Cx<T>>();
      Y.M2 = M2;
      HIPC(launch_spec<T>(spec_[0].n, spec_[0].grid, spec_[0].block, spec_[0].shmem, stream_, Y, spec_[0].pv));
    }
    C2RArgs<T> c = ca_;
    c.S2 = nullptr;
    c.M2 = nullptr;
    c.p = Q;
    c.S1 = S1_.as<Cx<T>>();
    if (nd_ == 1) {
      // 1-D: the multipliers are applied in the c2r pass
      HIPC(hipMemcpyAsync(S2_.p, S1_.p, (size_t)NHP_ * sizeof(Cx<T>), hipMemcpyDeviceToDevice, stream_));
      HIPC(launch_c2r<T>(c2r_.n, c2r_.grid, c2r_.block, c2r_.shmem, stream_, c, c2r_.pv));
      c.S1 = S2_.as<Cx<T>>();
      c.M1 = M2;
      c.p = Qd;
      HIPC(launch_c2r<T>(c2r_.n, c2r_.grid, c2r_.block, c2r_.shmem, stream_, c, c2r_.pv));
      return;
    }
    HIPC(launch_c2r<T>(c2r_.n, c2r_.grid, c2r_.block, c2r_.shmem, stream_, c, c2r_.pv));
    c.S1 = S2_.as<Cx<T>>();
    c.p = Qd;
    HIPC(launch_c2r<T>(c2r_.n, c2r_.grid, c2r_.block, c2r_.shmem, stream_, c, c2r_.pv));
  }

  // d/dc_ref [kappa (1 + nabla1)] / N on the octant (heterogeneous branch)
  void build_dkappa(DevBuf& out) {
    const long long hx = N_[0] / 2 + 1, hy = N_[1] / 2 + 1, hz = nd_ > 2 ? N_[2] / 2 + 1 : 1;
    std::vector<double> kx = wavenumbers(N_[0], d_[0]);
    std::vector<double> ky = nd_ > 1 ? wavenumbers(N_[1], d_[1]) : std::vector<double>{0.0};
    std::vector<double> kz = nd_ > 2 ? wavenumbers(N_[2], d_[2]) : std::vector<double>{0.0};
    std::vector<T> m((size_t)NHP_ * hy * hz, (T)0);   // rows padded to NHP_ like M1
    const double inv_n = 1.0 / (double)plane_;
    for (long long iz = 0; iz < hz; ++iz)
      for (long long iy = 0; iy < hy; ++iy)
        for (long long ix = 0; ix < hx; ++ix) {
          const double kk = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz]);
          const double a = c_ref_ * kk * dt_ / 2;
          const double kap = a == 0.0 ? 1.0 : std::sin(a) / a;
          const double dkap = a == 0.0 ? 0.0 : (a * std::cos(a) - std::sin(a)) / (a * a) * (kk * dt_ / 2);
          double n1 = 0.0, dn1 = 0.0;
          if (absorbing_) {
            n1 = tau_ * std::pow(kk, y_);
            dn1 = -2.0 * anp_ * (y_ - 1.0) * std::pow(c_ref_, y_ - 2.0) * std::pow(kk, y_);
          }
          m[ix + NHP_ * (iz + hz * iy)] = (T)((dkap * (1.0 + n1) + kap * dn1) * inv_n);
        }
    upload(out, m);
  }

  template <typename U>
  friend class Batch;
  enum Cls { C_DIR_X, C_DIR_Y, C_DIR_Z, C_SPEC_Y_FWD, C_SPEC_Z, C_SPEC_Y_INV, C_SPEC_Y, C_C2R, C_REC, C_TR,
             C_PLANE_C2R, kNumClasses };
  static constexpr const char* kClassNames[kNumClasses] = {
      "k_dir_pass_x_eos", "k_dir_pass_y", "k_dir_pass_z", "k_spec_pass_y_fwd", "k_spec_pass_z_mul",
      "k_spec_pass_y_inv", "k_spec_pass_y_mul", "k_c2r_x", "k_record", "k_tr_inject", "k_plane_c2r"};
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
  };
  struct Prof {
    std::vector<ProfRec> ev;
  };

  template <typename F>
  void launch(Prof* pr, int cls, F&& f) {
    if (!pr) {
      f();
      return;
    }
    ProfRec r;
    r.cls = cls;
    HIPC(hipEventCreate(&r.a));
    HIPC(hipEventCreate(&r.b));
    HIPC(hipEventRecord(r.a, stream_));
    f();
    HIPC(hipEventRecord(r.b, stream_));
    pr->ev.push_back(r);
  }

  void copy_out(const T* dsrc, long long n, double* out) {
    std::vector<T> h(n);
    if (n) HIPC(hipMemcpy(h.data(), dsrc, n * sizeof(T), hipMemcpyDeviceToHost));
    for (long long i = 0; i < n; ++i) out[i] = (double)h[i];
  }

  // ---- construction -------------------------------------------------------
  void build(const kw_sim_desc& d) {
    HIPC(hipSetDevice(dev_));
    nd_ = d.ndim;
    if (nd_ < 1 || nd_ > 3) raise(KW_ERR_INVALID, "ndim must be 1, 2 or 3");
    for (int a = 0; a < 3; ++a) N_[a] = a < nd_ ? d.N[a] : 1;
    if (nfire_ < 1) raise(KW_ERR_INVALID, "firing count must be positive");
    if (nfire_ > 1) {
      if (nd_ != 2) raise(KW_ERR_UNSUPPORTED, "stacked firings need 2-D grids");
      N_[2] = nfire_;
    }
    for (int a = 0; a < nd_; ++a) {
      if (N_[a] < 1) raise(KW_ERR_INVALID, "grid size must be positive");
      // compiled plan, or a chirp-z (Bluestein) line on a power-of-two plan
      bs_[a] = !size_supported(N_[a]);
      M_[a] = bs_[a] ? bluestein_plan(N_[a]) : N_[a];
      if (M_[a] == 0)
        raise(KW_ERR_UNSUPPORTED, "axis length " + std::to_string(N_[a]) +
                                      " is longer than the largest chirp-z line (2048) and has no compiled plan");
      if (bs_[a] && batch_mode_)
        raise(KW_ERR_UNSUPPORTED, "axis length " + std::to_string(N_[a]) +
                                      " has no compiled plan; batches take compiled lengths only");
      if (d.pml_size[a] < 0) raise(KW_ERR_INVALID, "pml_size must be >= 0");
      if (2LL * d.pml_size[a] > N_[a] + 0 && d.pml_size[a] > N_[a])
        raise(KW_ERR_INVALID, "pml_size larger than the grid");
    }
    if (d.dt == 0.0 || d.Nt == 0)
      raise(KW_ERR_INVALID, "Time array not set. Call make_time!(kgrid, sound_speed) before running.");
    check_lane_offsets(nd_, N_, (int)sizeof(T));
    if (d.map_storage != KW_MAPS_NATIVE && d.map_storage != KW_MAPS_FP16)
      raise(KW_ERR_INVALID, "unknown map storage mode");
    maps_f16_ = d.map_storage == KW_MAPS_FP16;
    {
      int v = 0;
      HIPC(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_));
      lds_max_ = v > 0 ? (size_t)v : 65536;
      int cus = 0;
      HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_));
      fill_wgs_ = 2LL * std::max(cus, 1);
    }
    nvox_ = N_[0] * N_[1] * N_[2];
    plane_ = nvox_ / nfire_;
    NH_ = N_[0] / 2 + 1;
    NHP_ = (NH_ + 15) / 16 * 16;  // 128-byte aligned spectrum rows (FP32)
    dt_ = d.dt;
    Nt_ = d.Nt;
    Nt_alloc_ = d.Nt;
    axisym_ = d.axisymmetric != 0;
    if (axisym_ && nd_ != 2) raise(KW_ERR_INVALID, "the axisymmetric solver needs a 2-D grid (axial x, radial y)");
    tr_ = d.tr_data != nullptr && !axisym_;   // axisymmetric.jl has no time-reversal branch
    HIPC(hipStreamCreate(&stream_));

    const bool het_c = d.c0 != nullptr, het_r = d.rho0 != nullptr;
    het_c_ = het_c;
    nonlinear_ = d.nonlinear != 0;
    has_p0_ = d.p0 != nullptr && !tr_;
    // fields
    p_.alloc(nvox_ * sizeof(T));
    HIPC(hipMemset(p_.p, 0, p_.bytes));
    for (int a = 0; a < nd_; ++a) {
      u_[a].alloc(nvox_ * sizeof(T));
      rho_[a].alloc(nvox_ * sizeof(T));
      HIPC(hipMemset(u_[a].p, 0, u_[a].bytes));
      HIPC(hipMemset(rho_[a].p, 0, rho_[a].bytes));
    }
    // medium maps: T, or FP16 offsets from c0_offset / rho0_offset
    if (maps_f16_) {
      if (het_r) upload(rho0_, to_half(d.rho0, d.rho0_offset));
      if (het_c) upload(c0sq_, to_half(d.c0, d.c0_offset));
    } else {
      if (het_r) upload(rho0_, to_t<T>(std::vector<double>(d.rho0, d.rho0 + nvox_)));
      if (het_c) {
        std::vector<double> c2(nvox_);
        for (long long i = 0; i < nvox_; ++i) c2[i] = d.c0[i] * d.c0[i];
        upload(c0sq_, to_t<T>(c2));
      }
    }
    if (d.nonlinear && d.bona) upload(bona_, to_t<T>(std::vector<double>(d.bona, d.bona + nvox_)));

    // axis tables
    for (int a = 0; a < nd_; ++a) {
      AxisTables A = axis_tables(N_[a], d.d[a], d.pml_size[a], d.pml_alpha[a], d.dt);
      if (bs_[a]) {
        // the line FFT runs on the M-point plan; chirp tables carry length N
        std::vector<double> wr, wi;
        twiddles(M_[a], wr, wi);
        upload(tw_[a], to_cx<T>(wr, wi));
        const ChirpTables C = chirp_tables(N_[a], M_[a]);
        upload(chc_[a], to_cx<T>(C.cr, C.ci));
        upload(chf_[a], to_cx<T>(C.gfr, C.gfi));
        upload(chb_[a], to_cx<T>(C.gbr, C.gbi));
      } else {
        upload(tw_[a], to_cx<T>(A.twr, A.twi));
      }
      upload(Dp_[a], to_cx<T>(A.dpr, A.dpi));
      upload(Dm_[a], to_cx<T>(A.dmr, A.dmi));
      upload(pml_[a], to_t<T>(A.pml));
      upload(pmlsg_[a], to_t<T>(A.pml_sg));
    }
    if (nd_ > 1) {
      // the x half spectrum is transformed along y/z with the full-length twiddles
    }
    build_eos(d, het_c);

    // launch geometry

    // launch shapes follow the plan length (M_ = N_ unless the axis runs chirp-z lines)
    cfg_[0] = pass_cfg(M_[0], true, N_[1] * N_[2], 0, TC_DIR_X, bs_[0]);
    if (nd_ > 1) cfg_[1] = pass_cfg(M_[1], false, N_[0], N_[2], TC_DIR_Y);
    if (nd_ > 2) cfg_[2] = pass_cfg(M_[2], false, N_[0], N_[1], TC_DIR_Z);
    if (nd_ > 1) spec_[0] = spec_cfg(M_[1], N_[2], TC_SPEC_Y);
    if (nd_ > 2) spec_[1] = spec_cfg(M_[2], N_[1], TC_SPEC_Z);
    c2r_ = pass_cfg(M_[0], true, N_[1] * N_[2], 0, TC_C2R, bs_[0]);

    build_sources(d, het_c);
    build_sensor(d);
    build_args(d);
    if (d.p0 && !tr_) init_p0(d);
    HIPC(hipStreamSynchronize(stream_));
  }

  // IEEE binary16 bits of (v - off), rounded to nearest even (clang _Float16)
  std::vector<uint16_t> to_half(const double* v, double off) {
    std::vector<uint16_t> h(nvox_);
    for (long long i = 0; i < nvox_; ++i) {
      const _Float16 x = (_Float16)(v[i] - off);
      std::memcpy(&h[i], &x, sizeof(uint16_t));
    }
    return h;
  }

  PassCfg pass_cfg(long long n, bool contig, long long lines, long long outer, int cls, bool bs = false) {
    PassCfg c;
    c.n = (int)n;
    // plan family: 0 strided, 1 contiguous, 2 complex-to-real, 4 contiguous chirp-z lines, or a tuned
    // alternative (single simulations on compiled lengths; batch kernels take the defaults)
    const int dflt_v = contig && bs ? 4 : (cls == TC_C2R ? 2 : (contig ? 1 : 0));
    c.pv = (bs || batch_mode_) ? -1 : tuned_plan((int)sizeof(T), cls, (int)n);
    if (c.pv == dflt_v) c.pv = -1;   // the default family: its own kernels
    const int tpl = plan_tpl((int)n, c.pv >= 0 ? c.pv : dflt_v);
    const size_t line_bytes = (size_t)plan_pitch((int)n) * sizeof(Cx<T>);
    // strided lines: at least 16 line pairs per workgroup (128-byte rows in FP32)
    // when that fits 512 threads; contiguous lines: 256 threads
    int lpw = (contig ? 256 : std::max(256, std::min(512, 16 * tpl))) / tpl;
    while (lpw > 1 && (size_t)lpw * line_bytes > lds_max_) lpw /= 2;
    int maxl = (cls == TC_C2R ? 256 : 512) / tpl;   // the kernels' launch bounds
    while (maxl > 1 && (size_t)maxl * line_bytes > lds_max_) maxl /= 2;
    lpw = tuned_lpw((int)sizeof(T), cls, (int)n, lpw, std::max(maxl, 1));
    c.req = lpw;
    // pairs that exist
    const long long pairs = (lines + 1) / 2;
    while (!batch_mode_ && lpw > 1 && lpw / 2 >= pairs) lpw /= 2;
    // small grids (2-D, few lines): fewer lines per workgroup until the launch
    // has two workgroups per CU, keeping at least one full wave per workgroup
    if (!batch_mode_) {
      const int min_lpw = std::max(1, 64 / tpl);
      const long long o = contig ? 1 : std::max<long long>(outer, 1);
      while (lpw / 2 >= min_lpw && ((pairs + lpw - 1) / lpw) * o < fill_wgs_) lpw /= 2;
    }
    c.lpw = lpw;
    c.block = dim3(lpw * tpl);
    c.shmem = (size_t)lpw * line_bytes;
    if (c.shmem > lds_max_)
      raise(KW_ERR_UNSUPPORTED, "line of length " + std::to_string(n) + " exceeds the LDS of a workgroup");
    if (contig) {
      c.grid = dim3((unsigned)((pairs + lpw - 1) / lpw));
    } else {
      const long long xb = (pairs + lpw - 1) / lpw;
      c.grid = dim3((unsigned)(xb * std::max<long long>(outer, 1)));
    }
    return c;
  }
  PassCfg spec_cfg(long long n, long long outer, int cls) {
    PassCfg c;
    c.n = (int)n;
    c.pv = (bs_[cls == TC_SPEC_Y ? 1 : 2] || batch_mode_) ? -1 : tuned_plan((int)sizeof(T), cls, (int)n);
    if (c.pv == 0) c.pv = -1;   // the default family: its own kernels
    const int tpl = plan_tpl((int)n, c.pv >= 0 ? c.pv : 0);
    const size_t line_bytes = (size_t)plan_pitch((int)n) * sizeof(Cx<T>);
    int lpw = std::max(256, std::min(512, 16 * tpl)) / tpl;   // as for strided passes
    while (lpw > 1 && (size_t)lpw * line_bytes > lds_max_) lpw /= 2;
    int maxl = 512 / tpl;
    while (maxl > 1 && (size_t)maxl * line_bytes > lds_max_) maxl /= 2;
    lpw = tuned_lpw((int)sizeof(T), cls, (int)n, lpw, std::max(maxl, 1));
    c.req = lpw;
    while (!batch_mode_ && lpw > 1 && lpw / 2 >= NH_) lpw /= 2;
    if (!batch_mode_) {   // as in pass_cfg: fill the device on small grids
      const int min_lpw = std::max(1, 64 / tpl);
      while (lpw / 2 >= min_lpw && ((NH_ + lpw - 1) / lpw) * std::max<long long>(outer, 1) < fill_wgs_) lpw /= 2;
    }
    c.lpw = lpw;
    c.block = dim3(lpw * tpl);
    c.shmem = (size_t)lpw * line_bytes;
    const long long xb = (NH_ + lpw - 1) / lpw;
    c.grid = dim3((unsigned)(xb * std::max<long long>(outer, 1)));
    return c;
  }

  // Equation-of-state multipliers (first_order.jl:13-42, 156-170;
  // first_order_steps.jl:214-290) on the octant |kx| x |ky| x |kz| with the
  // inverse-transform 1/(Nx Ny Nz) folded in.
  void build_eos(const kw_sim_desc& d, bool het_c) {
    const long long Nt3[3] = {N_[0], N_[1], nd_ > 2 ? N_[2] : 1};   // one firing's grid
    const EosTables E = eos_tables(d, nd_, Nt3, het_c);
    const double c_ref = d.c_ref;
    const double inv_n = 1.0 / (double)plane_;
    stokes_ = E.stokes;
    c_ref_ = c_ref;
    absorbing_ = E.absorbing;
    tau_ = E.tau;
    y_ = d.alpha_power;
    anp_ = E.absorbing ? -E.tau / (2.0 * std::pow(c_ref, y_ - 1.0)) : 0.0;
    for (int q = 0; q < 3; ++q) d_[q] = d.d[q];
    // device rows are padded to the spectrum's row pitch (NHP_, a multiple of 16
    // entries) so a workgroup's run of 16 kx multipliers is one aligned segment
    const size_t rows = (size_t)(E.hy * E.hz);
    std::vector<T> m1((size_t)NHP_ * rows, (T)0), m2(E.stokes ? m1.size() : 0, (T)0);
    for (size_t r = 0; r < rows; ++r)
      for (long long ix = 0; ix < E.hx; ++ix) {
        m1[r * NHP_ + ix] = (T)(E.m1[r * E.hx + ix] * inv_n);
        if (E.stokes) m2[r * NHP_ + ix] = (T)(E.m2[r * E.hx + ix] * inv_n);
      }
    upload(M1_, m1);
    if (stokes_) upload(M2_, m2);
    // spectra
    const long long lines = N_[1] * N_[2];
    S1_.alloc((size_t)NHP_ * lines * sizeof(Cx<T>));
    if (stokes_) {
      S2_.alloc((size_t)NHP_ * lines * sizeof(Cx<T>));
      rho_prev_.alloc(nvox_ * sizeof(T));
      HIPC(hipMemset(rho_prev_.p, 0, rho_prev_.bytes));
    }
  }

  // Per-workgroup source entry lists for each directional pass.
  // psrc: density injection p/(ndim c^2) (first_order_steps.jl:561-580,
  // 619-640, 692-715); usrc: raw velocity values (541-559, 586-617, 646-690).
  void build_sources(const kw_sim_desc& d, bool het_c) {
    if (d.n_psrc > 0 && d.psrc_index && d.psrc_signal) {
      const long long ns = d.n_psrc, nc = std::max<long long>(d.psrc_ncols, 1);
      std::vector<T> vals(ns * nc);
      for (long long j = 0; j < ns; ++j) {
        const long long idx = d.psrc_index[j];
        if (idx < 0 || idx >= nvox_) raise(KW_ERR_INVALID, "pressure source index out of range");
        const double c = het_c ? d.c0[idx] : d.c0_scalar;
        const double sc = (double)nd_ * c * c;
        for (long long k = 0; k < nc; ++k) vals[k * ns + j] = (T)(d.psrc_signal[j + ns * k] / sc);
      }
      upload(pval_, vals);
      p_ns_ = ns;
      p_nc_ = nc;
      psrc_idx_h_.assign(d.psrc_index, d.psrc_index + ns);
      psrc_sig_h_.assign(d.psrc_signal, d.psrc_signal + ns * nc);
      psrc_fac_h_.resize(ns);
      for (long long j = 0; j < ns; ++j) {
        const double c = het_c ? d.c0[d.psrc_index[j]] : d.c0_scalar;
        psrc_fac_h_[j] = -2.0 / ((double)nd_ * c * c * c);   // d/dc of 1/(ndim c^2)
      }
      p_dir_ = d.p_mode == KW_DIRICHLET;
      for (int a = 0; a < nd_; ++a) build_csr(a, d.psrc_index, ns, psrc_[a]);
    }
    if (d.n_usrc > 0 && d.usrc_index) {
      for (int a = 0; a < nd_; ++a) {
        if (!d.usrc_signal[a]) continue;
        const long long ns = d.n_usrc, nc = std::max<long long>(d.usrc_ncols[a], 1);
        std::vector<T> vals(ns * nc);
        for (long long j = 0; j < ns; ++j) {
          if (d.usrc_index[j] < 0 || d.usrc_index[j] >= nvox_)
            raise(KW_ERR_INVALID, "velocity source index out of range");
          for (long long k = 0; k < nc; ++k) vals[k * ns + j] = (T)d.usrc_signal[a][j + ns * k];
        }
        upload(uval_[a], vals);
        u_ns_ = ns;
        u_nc_[a] = nc;
        build_csr(a, d.usrc_index, ns, usrc_[a]);
      }
      u_dir_ = d.u_mode == KW_DIRICHLET;
    }
  }

  void build_csr(int axis, const int64_t* index, long long ns, SrcDev& out) {
    const PassCfg& c = cfg_[axis];
    const long long nwg = c.grid.x;
    std::vector<long long> wg(ns);
    std::vector<int> pos(ns);
    const long long Nx = N_[0], Ny = N_[1];
    const long long xblocks = ((Nx + 1) / 2 + c.lpw - 1) / c.lpw;
    for (long long j = 0; j < ns; ++j) {
      const long long idx = index[j];
      const long long x = idx % Nx, y = (idx / Nx) % Ny, z = idx / (Nx * Ny);
      long long w, cc, part, n;
      if (axis == 0) {
        const long long L = idx / Nx;
        w = (L / 2) / c.lpw;
        cc = (L / 2) % c.lpw;
        part = L % 2;
        n = x;
      } else {
        const long long xb = (x / 2) / c.lpw;
        cc = (x / 2) % c.lpw;
        part = x % 2;
        const long long o = axis == 1 ? z : y;
        n = axis == 1 ? y : z;
        w = xb + xblocks * o;
      }
      wg[j] = w;
      pos[j] = (int)(n | (part << 20) | (cc << 21));
    }
    std::vector<long long> order(ns);
    for (long long j = 0; j < ns; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](long long a, long long b) { return wg[a] < wg[b]; });
    std::vector<int> off(nwg + 1, 0), p(ns), r(ns);
    for (long long j = 0; j < ns; ++j) off[wg[j] + 1]++;
    for (long long w = 0; w < nwg; ++w) off[w + 1] += off[w];
    for (long long k = 0; k < ns; ++k) {
      p[k] = pos[order[k]];
      r[k] = (int)order[k];
    }
    upload(out.off, off);
    upload(out.pos, p);
    upload(out.row, r);
    out.any = true;
  }

  void build_sensor(const kw_sim_desc& d) {
    nsens_ = d.n_sensor;
    rec_ = d.record;
    dir_ = d.directivity != 0;
    fr_ = d.freq_response != 0;
    if (dir_) {
      dir_size_ = d.directivity_size;
      dir_ang_s_ = d.directivity_angle_scalar;
      if (d.directivity_angle && d.n_sensor > 0)
        upload(dir_ang_, std::vector<double>(d.directivity_angle, d.directivity_angle + d.n_sensor));
    }
    if (fr_) {
      fc_ = d.freq_center;
      bw_ = d.freq_bandwidth;
    }
    dmin_ = nd_ == 1 ? d.d[0] : *std::min_element(d.d, d.d + nd_);
    cartesian_ = d.n_sensor > 0 && d.sensor_index == nullptr;
    if (tr_) {
      // TR injection values p_TR[j, t] / (ndim c^2) (first_order.jl:234-242, 450-460, 678-689)
      if (d.n_sensor > 0 && d.sensor_index) {
        const long long n = d.n_sensor;
        std::vector<T> vals(n * Nt_);
        for (long long j = 0; j < n; ++j) {
          const long long idx = d.sensor_index[j];
          const double c = d.c0 ? d.c0[idx] : d.c0_scalar;
          const double sc = (double)nd_ * c * c;
          for (long long k = 0; k < Nt_; ++k) vals[k * n + j] = (T)(d.tr_data[j + n * k] / sc);
        }
        upload(tr_vals_, vals);
      }
    }
    if (nsens_ == 0 || cartesian_) return;
    std::vector<long long> idx(d.sensor_index, d.sensor_index + nsens_);
    for (long long v : idx)
      if (v < 0 || v >= nvox_) raise(KW_ERR_INVALID, "sensor index out of range");
    upload(sidx_, idx);
    // a mask covering the whole grid records voxel j in row j: the recorder skips the index stream
    dense_sensor_ = nsens_ == nvox_;
    for (long long j = 0; j < nsens_ && dense_sensor_; ++j) dense_sensor_ = idx[j] == j;
    auto fill_t = [&](DevBuf& b, long long n, double v) {
      upload(b, std::vector<T>(n, (T)v));
    };
    auto fill_d = [&](DevBuf& b, long long n, double v) { upload(b, std::vector<double>(n, v)); };
    if (rec_ & KW_REC_P) {
      p_ts_.alloc((size_t)nsens_ * Nt_alloc_ * sizeof(T));
      HIPC(hipMemset(p_ts_.p, 0, p_ts_.bytes));
    }
    for (int a = 0; a < 3; ++a) {
      const uint32_t bit = a == 0 ? KW_REC_UX : (a == 1 ? KW_REC_UY : KW_REC_UZ);
      if ((rec_ & bit) && a < nd_) {
        u_ts_[a].alloc((size_t)nsens_ * Nt_alloc_ * sizeof(T));
        HIPC(hipMemset(u_ts_[a].p, 0, u_ts_[a].bytes));
      }
    }
    if (rec_ & KW_REC_P_MAX) fill_t(p_max_, nsens_, -INFINITY);
    if (rec_ & KW_REC_P_MIN) fill_t(p_min_, nsens_, INFINITY);
    if (rec_ & KW_REC_P_RMS) fill_d(p_rms_, nsens_, 0.0);
    if (rec_ & KW_REC_P_FINAL) fill_t(p_final_, nsens_, 0.0);
    if (rec_ & KW_REC_U_MAX) fill_d(u_max_, nsens_, -INFINITY);
    if (rec_ & KW_REC_U_RMS) fill_d(u_rms_, nsens_, 0.0);
    if (rec_ & KW_REC_U_FINAL) fill_d(u_final_, nsens_, 0.0);
    if (rec_ & KW_REC_I_AVG) fill_d(I_avg_, nsens_, 0.0);
    if (rec_ & KW_REC_I_MAX) fill_d(I_max_, nsens_, -INFINITY);
    if ((rec_ & (KW_REC_I_AVG | KW_REC_I_MAX)) && sizeof(T) != 8)
      raise(KW_ERR_INVALID,
            "I_avg/I_max require data_cast=Float64 (_acoustic_intensity is typed for Float64)");
  }

  void build_args(const kw_sim_desc& d) {
    const bool het_r = d.rho0 != nullptr;
    if (axisym_) {
      // r = |y_vec|, y_vec = (j - (Ny-1)/2) dy, 0 replaced by dy/2 (axisymmetric.jl:103-106)
      std::vector<double> ir(N_[1]);
      for (long long j = 0; j < N_[1]; ++j) {
        double r = std::fabs(((double)j - (double)(N_[1] - 1) / 2.0) * d.d[1]);
        if (r == 0.0) r = d.d[1] / 2.0;
        ir[j] = 1.0 / r;
      }
      upload(inv_r_, to_t<T>(ir));
    }
    for (int a = 0; a < nd_; ++a) {
      PassArgs<T>& A = pa_[a];
      std::memset(&A, 0, sizeof(A));
      A.p = p_.as<T>();
      A.u = u_[a].as<T>();
      A.rho = rho_[a].as<T>();
      A.rho0 = het_r ? rho0_.p : nullptr;
      A.rho0_s = (T)d.rho0_scalar;
      A.maps_f16 = maps_f16_ ? 1 : 0;
      A.rho0_off = (T)d.rho0_offset;
      A.c0_off = (T)d.c0_offset;
      A.pml = pml_[a].as<T>();
      A.pml_sg = pmlsg_[a].as<T>();
      A.Dp = Dp_[a].as<Cx<T>>();
      A.Dm = Dm_[a].as<Cx<T>>();
      A.tw = tw_[a].as<Cx<T>>();
      A.dt = (T)d.dt;
      A.Nx = (int)N_[0];
      A.n_lines = (int)(N_[1] * N_[2]);
      A.inv_r = (axisym_ && a == 1) ? inv_r_.as<T>() : nullptr;
      A.nl = (int)N_[a];
      A.ch = chirp(a);
      const PassCfg& c = cfg_[a];
      A.x_blocks = (int)(((N_[0] + 1) / 2 + c.lpw - 1) / c.lpw);
      if (a == 1) {
        A.Sa = N_[0];
        A.So = N_[0] * N_[1];
        A.n_outer = (int)N_[2];
      } else if (a == 2) {
        A.Sa = N_[0] * N_[1];
        A.So = N_[0];
        A.n_outer = (int)N_[1];
      } else {
        A.Sa = 1;
      }
      if (psrc_[a].any) {
        A.psrc = SrcList{psrc_[a].off.as<int>(), psrc_[a].pos.as<int>(), psrc_[a].row.as<int>(), pval_.p,
                         (int)p_ns_, (int)p_nc_, p_dir_ ? 1 : 0};
      }
      if (usrc_[a].any) {
        A.usrc = SrcList{usrc_[a].off.as<int>(), usrc_[a].pos.as<int>(), usrc_[a].row.as<int>(), uval_[a].p,
                         (int)u_ns_, (int)u_nc_[a], u_dir_ ? 1 : 0};
      }
    }
    // the x pass carries the EOS input and its r2c stage
    PassArgs<T>& X = pa_[0];
    X.eos = 1;
    X.rho_o1 = nd_ > 1 ? rho_[1].as<T>() : nullptr;
    X.rho_o2 = nd_ > 2 ? rho_[2].as<T>() : nullptr;
    X.c0sq = c0sq_.p;
    X.nonlinear = d.nonlinear ? 1 : 0;
    X.bona = bona_.p ? bona_.as<T>() : nullptr;
    X.bona_s = (T)d.bona_scalar;
    X.S1 = S1_.as<Cx<T>>();
    X.S2 = stokes_ ? S2_.as<Cx<T>>() : nullptr;
    X.rho_prev = stokes_ ? rho_prev_.as<T>() : nullptr;
    X.inv_dt = (T)(1.0 / d.dt);
    X.nhp = (int)NHP_;

    // spectral passes
    const int hy = (int)(N_[1] / 2 + 1), hz = (int)(N_[2] / 2 + 1);
    (void)hz;
    if (nd_ > 1) {
      SpecArgs<T>& Y = sa_[0];
      std::memset(&Y, 0, sizeof(Y));
      Y.S1 = S1_.as<Cx<T>>();
      Y.S2 = stokes_ ? S2_.as<Cx<T>>() : nullptr;
      Y.tw = tw_[1].as<Cx<T>>();
      Y.M1 = M1_.as<T>();
      Y.M2 = stokes_ ? M2_.as<T>() : nullptr;
      Y.NH = (int)NH_;
      Y.Sc = 1;
      Y.Sa = NHP_;
      Y.So = NHP_ * N_[1];
      Y.x_blocks = (int)((NH_ + spec_[0].lpw - 1) / spec_[0].lpw);
      Y.axis = 1;
      Y.Ny = (int)N_[1];
      Y.Nz = nd_ > 2 ? (int)N_[2] : 1;   // stacked firings: no octant index along the firing axis
      Y.nl = (int)N_[1];
      Y.ch = chirp(1);
      Y.MH1 = (int)NHP_;   // padded table row pitch
      Y.MH2 = hy;
    }
    if (nd_ > 2) {
      SpecArgs<T>& Z = sa_[1];
      std::memset(&Z, 0, sizeof(Z));
      Z.S1 = S1_.as<Cx<T>>();
      Z.S2 = stokes_ ? S2_.as<Cx<T>>() : nullptr;
      Z.tw = tw_[2].as<Cx<T>>();
      Z.M1 = M1_.as<T>();
      Z.M2 = stokes_ ? M2_.as<T>() : nullptr;
      Z.NH = (int)NH_;
      Z.Sc = 1;
      Z.Sa = NHP_ * N_[1];
      Z.So = NHP_;
      Z.x_blocks = (int)((NH_ + spec_[1].lpw - 1) / spec_[1].lpw);
      Z.axis = 2;
      Z.Ny = (int)N_[1];
      Z.Nz = (int)N_[2];
      Z.nl = (int)N_[2];
      Z.ch = chirp(2);
      Z.MH1 = (int)NHP_;
      Z.MH2 = hz;
    }
    std::memset(&ca_, 0, sizeof(ca_));
    ca_.S1 = S1_.as<Cx<T>>();
    ca_.S2 = (nd_ == 1 && stokes_) ? S2_.as<Cx<T>>() : nullptr;
    ca_.M1 = nd_ == 1 ? M1_.as<T>() : nullptr;
    ca_.M2 = (nd_ == 1 && stokes_) ? M2_.as<T>() : nullptr;
    ca_.tw = tw_[0].as<Cx<T>>();
    ca_.p = p_.as<T>();
    ca_.n_lines = (int)(N_[1] * N_[2]);
    // the fused y-inverse / c2r plane pass where the plane fits a workgroup
    plane_c2r_ = nd_ == 3 && !batch_mode_ && !bs_[0] && !bs_[1] && plane_c2r_supported<T>((int)N_[0], (int)N_[1], lds_max_);
    std::memset(&plane_args_, 0, sizeof(plane_args_));
    plane_args_.S1 = S1_.as<Cx<T>>();
    plane_args_.tw_x = tw_[0].as<Cx<T>>();
    plane_args_.tw_y = nd_ > 1 ? tw_[1].as<Cx<T>>() : nullptr;
    plane_args_.p = p_.as<T>();
    plane_args_.nhp = (int)NHP_;
    ca_.nhp = (int)NHP_;
    ca_.nl = (int)N_[0];
    ca_.ch = chirp(0);

    // recording
    std::memset(&ra_, 0, sizeof(ra_));
    ra_.p = p_.as<T>();
    for (int a = 0; a < 3; ++a) ra_.u[a] = a < nd_ ? u_[a].as<T>() : nullptr;
    ra_.ndim = nd_;
    ra_.idx = dense_sensor_ ? nullptr : sidx_.as<long long>();
    ra_.n = (int)nsens_;
    ra_.Nt = (int)Nt_;
    ra_.rec = rec_;
    ra_.p_ts = p_ts_.as<T>();
    for (int a = 0; a < 3; ++a) ra_.u_ts[a] = u_ts_[a].as<T>();
    ra_.p_max = p_max_.as<T>();
    ra_.p_min = p_min_.as<T>();
    ra_.p_rms = p_rms_.as<double>();
    ra_.p_final = p_final_.as<T>();
    ra_.u_max = u_max_.as<double>();
    ra_.u_rms = u_rms_.as<double>();
    ra_.u_final = u_final_.as<double>();
    ra_.I_avg = I_avg_.as<double>();
    ra_.I_max = I_max_.as<double>();
  }

  // initialize_p0_{1,2,3}d! (first_order_steps.jl:58-191) with p0 already smoothed
  void init_p0(const kw_sim_desc& d) {
    DevBuf& p0 = p0_keep_;
    upload(p0, to_t<T>(std::vector<double>(d.p0, d.p0 + nvox_)));
    if (d.smooth_p0) smooth_p0(p0);
    // rho_d = p0 / (ndim c0^2); p = p0
    const unsigned nb = (unsigned)((nvox_ + 255) / 256);
    DevBuf c2tmp;
    if (d.c0) {
      std::vector<double> c2(nvox_);
      for (long long i = 0; i < nvox_; ++i) c2[i] = d.c0[i] * d.c0[i];
      upload(c2tmp, to_t<T>(c2));
    }
    HIPC(launch_init_rho<T>(nb, stream_, p0.as<T>(), d.c0 ? c2tmp.as<T>() : nullptr,
                            (T)(d.c0_scalar * d.c0_scalar), (T)nd_, p_.as<T>(), rho_[0].as<T>(),
                            nd_ > 1 ? rho_[1].as<T>() : nullptr, nd_ > 2 ? rho_[2].as<T>() : nullptr, nvox_));
    // u_d = -dt/(2 rho0_sg) d/dd p0
    for (int a = 0; a < nd_; ++a) {
      PassArgs<T> A = pa_[a];
      A.mode = 1;
      A.eos = 0;
      HIPC(ldir(a, cfg_[a], A));
    }
    HIPC(hipStreamSynchronize(stream_));
  }

  // smooth(p0; restore_max=true) (filters.jl:20-60): per axis FFT, multiply
  // by fftshift(blackman(N)), IFFT on complex data; then rescale so that
  // max|smoothed| = max|p0|.
  void smooth_p0(DevBuf& p0) {
    DevBuf z, mx;
    z.alloc(nvox_ * sizeof(Cx<T>));
    mx.alloc(2 * sizeof(unsigned long long));
    HIPC(hipMemsetAsync(mx.p, 0, mx.bytes, stream_));
    const unsigned nb = (unsigned)((nvox_ + 255) / 256);
    HIPC(launch_smooth_helpers<T>(0, nb, stream_, p0.as<T>(), z.as<Cx<T>>(), nullptr, nullptr, nvox_));
    for (int a = 0; a < nd_; ++a) {
      const long long n = N_[a];
      std::vector<double> w(n), ws(n);
      for (long long i = 0; i < n; ++i)
        w[i] = n == 1 ? 1.0 : 0.42 - 0.5 * std::cos(2 * M_PI * i / (double)(n - 1)) +
                                  0.08 * std::cos(4 * M_PI * i / (double)(n - 1));
      for (long long i = 0; i < n; ++i) ws[i] = w[((i - n / 2) % n + n) % n] / (double)n;  // circshift(w, n/2)
      DevBuf wb;
      upload(wb, to_t<T>(ws));
      SpecArgs<T> A;
      std::memset(&A, 0, sizeof(A));
      A.S1 = z.as<Cx<T>>();
      A.tw = tw_[a].as<Cx<T>>();
      A.mode = 3;
      A.W = wb.as<T>();
      long long lines, outer;
      if (a == 0) {
        lines = N_[1] * N_[2];
        A.Sc = N_[0];
        A.Sa = 1;
        A.So = 0;
        outer = 1;
      } else {
        lines = N_[0];
        A.Sc = 1;
        A.Sa = a == 1 ? N_[0] : N_[0] * N_[1];
        A.So = a == 1 ? N_[0] * N_[1] : N_[0];
        outer = a == 1 ? N_[2] : N_[1];
      }
      A.NH = (int)lines;
      A.nl = (int)n;
      A.ch = chirp(a);
      PassCfg c = line_cfg(M_[a], lines, outer);
      A.x_blocks = (int)((lines + c.lpw - 1) / c.lpw);
      HIPC(lspec(a, c, A));
      HIPC(hipStreamSynchronize(stream_));
    }
    const unsigned nr = std::min<unsigned>(nb, 1024u);
    HIPC(launch_smooth_helpers<T>(1, nr, stream_, p0.as<T>(), z.as<Cx<T>>(), mx.as<unsigned long long>(), nullptr,
                                  nvox_));
    HIPC(launch_smooth_helpers<T>(2, nb, stream_, nullptr, z.as<Cx<T>>(), mx.as<unsigned long long>(), p0.as<T>(),
                                  nvox_));
    HIPC(hipStreamSynchronize(stream_));
  }

  // chirp tables of axis a (zero when its lines run a compiled plan)
  Chirp<T> chirp(int a) {
    Chirp<T> c{};
    if (bs_[a]) {
      c.n = (int)N_[a];
      c.c = chc_[a].as<Cx<T>>();
      c.gf = chf_[a].as<Cx<T>>();
      c.gb = chb_[a].as<Cx<T>>();
    }
    return c;
  }
  // pass launches on axis a: the compiled-length kernels or their chirp-z form
  hipError_t ldir(int a, const PassCfg& c, const PassArgs<T>& A) {
    return bs_[a] ? launch_dir_bs<T>(c.n, a == 0, c.grid, c.block, c.shmem, stream_, A)
                  : launch_dir<T>(c.n, a == 0, c.grid, c.block, c.shmem, stream_, A, c.pv);
  }
  hipError_t lspec(int a, const PassCfg& c, const SpecArgs<T>& A) {
    return bs_[a] ? launch_spec_bs<T>(c.n, c.grid, c.block, c.shmem, stream_, A)
                  : launch_spec<T>(c.n, c.grid, c.block, c.shmem, stream_, A, c.pv);
  }
  hipError_t lc2r(const PassCfg& c, const C2RArgs<T>& A) {
    return bs_[0] ? launch_c2r_bs<T>(c.n, c.grid, c.block, c.shmem, stream_, A)
                  : launch_c2r<T>(c.n, c.grid, c.block, c.shmem, stream_, A, c.pv);
  }

  // launch geometry for complex lines of length n (lines x outer)
  PassCfg line_cfg(long long n, long long lines, long long outer) {
    PassCfg c;
    c.n = (int)n;
    const int tpl = plan_tpl((int)n);
    const size_t line_bytes = (size_t)plan_pitch((int)n) * sizeof(Cx<T>);
    int lpw = 256 / tpl;
    while (lpw > 1 && (size_t)lpw * line_bytes > lds_max_) lpw /= 2;
    while (lpw > 1 && lpw / 2 >= lines) lpw /= 2;
    c.lpw = lpw;
    c.block = dim3(lpw * tpl);
    c.shmem = (size_t)lpw * line_bytes;
    c.grid = dim3((unsigned)(((lines + lpw - 1) / lpw) * std::max<long long>(outer, 1)));
    return c;
  }

  // ---- one time step ------------------------------------------------------
  void step(Prof* pr) {
    const int t = (int)(++t_);
    if (tr_ && tr_vals_.p && nsens_ > 0) {
      const long long n = nsens_;
      const T* v = tr_vals_.as<T>() + (Nt_ - t) * n;
      launch(pr, C_TR, [&] {
        HIPC(launch_tr_inject<T>((unsigned)((n + 255) / 256), stream_, rho_[0].as<T>(),
                                 nd_ > 1 ? rho_[1].as<T>() : nullptr, nd_ > 2 ? rho_[2].as<T>() : nullptr, sidx_tr(),
                                 v, (int)n));
      });
    }
    // y before z: whichever strided pass directly follows the previous step's
    // p writer runs ~4 us slower than in second place, and the z pass loses
    // more there than the y pass (A/B on one box: 0.4439 -> 0.4416 ms per
    // 256^3 FP32 step); the two passes write disjoint fields
    for (int a = 1; a < nd_; ++a) {
      pa_[a].t = t;
      launch(pr, a == 1 ? C_DIR_Y : C_DIR_Z,
             [&] { HIPC(ldir(a, cfg_[a], pa_[a])); });
    }
    pa_[0].t = t;
    launch(pr, C_DIR_X,
           [&] { HIPC(ldir(0, cfg_[0], pa_[0])); });
    if (nd_ == 3) {
      SpecArgs<T> Y = sa_[0];
      Y.mode = 0;
      launch(pr, C_SPEC_Y_FWD, [&] { HIPC(lspec(1, spec_[0], Y)); });
      SpecArgs<T> Z = sa_[1];
      Z.mode = 2;
      launch(pr, C_SPEC_Z, [&] { HIPC(lspec(2, spec_[1], Z)); });
      if (plane_c2r_) {   // y inverse and c2r fused per z plane
        launch(pr, C_PLANE_C2R, [&] {
          HIPC(launch_plane_c2r<T>((int)N_[0], (int)N_[1], (unsigned)N_[2], stream_, plane_args_));
        });
      } else {
        Y.mode = 1;
        Y.S2 = nullptr;
        launch(pr, C_SPEC_Y_INV, [&] { HIPC(lspec(1, spec_[0], Y)); });
      }
    } else if (nd_ == 2) {
      SpecArgs<T> Y = sa_[0];
      Y.mode = 2;
      launch(pr, C_SPEC_Y, [&] { HIPC(lspec(1, spec_[0], Y)); });
    }
    if (!plane_c2r_) launch(pr, C_C2R, [&] { HIPC(lc2r(c2r_, ca_)); });
    if (nsens_ > 0 && !cartesian_ && rec_ && !no_record_) {
      ra_.t = t;
      launch(pr, C_REC, [&] {
        HIPC(launch_record<T>((unsigned)((nsens_ + 255) / 256), stream_, ra_));
      });
    }
  }

  // apply_sensor_directivity! then apply_frequency_response! on the recorded
  // p (first_order_steps.jl:724-725, 836-910); computed once, in FP64.
  void apply_post_filters() {
    if (filtered_) return;
    const long long ns = nsens_, nt = Nt_alloc_;
    std::vector<double> cw(nt), sw(nt), absf(nt);
    for (long long m = 0; m < nt; ++m) {
      long long mm = m;
      double sg = -1.0;
      if (2 * mm > nt) {
        mm = nt - mm;
        sg = 1.0;
      }
      cw[m] = std::cos(2.0 * M_PI * (double)mm / (double)nt);
      sw[m] = sg * std::sin(2.0 * M_PI * (double)mm / (double)nt);
      // filters.jl:158-165: [0..N/2-1, -N/2..-1] (odd: symmetric) * Fs/N
      const long long k = (nt % 2 == 0) ? (m < nt / 2 ? m : m - nt) : (m <= (nt - 1) / 2 ? m : m - nt);
      absf[m] = std::fabs((double)k * ((1.0 / dt_) / (double)nt));
    }
    DevBuf dcw, dsw, dabs, Y;
    upload(dcw, cw);
    upload(dsw, sw);
    upload(dabs, absf);
    Y.alloc((size_t)ns * nt * sizeof(Cx<double>));
    pfilt_.alloc((size_t)ns * nt * sizeof(double));
    FiltArgs f;
    f.ns = (int)ns;
    f.nt = (int)nt;
    f.cw = dcw.as<double>();
    f.sw = dsw.as<double>();
    f.absf = dabs.as<double>();
    f.ang = dir_ang_.p ? dir_ang_.as<double>() : nullptr;
    f.ang_s = dir_ang_s_;
    f.dir = dir_ ? 1 : 0;
    f.dsize = dir_size_;
    f.kscale = 2.0 * M_PI / dmin_;
    f.fr = fr_ ? 1 : 0;
    f.fc = fc_;
    f.bw = bw_;
    f.Y = Y.as<Cx<double>>();
    f.out = pfilt_.as<double>();
    HIPC(launch_filter<T>(stream_, p_ts_.as<T>(), f));
    HIPC(hipStreamSynchronize(stream_));
    filtered_ = true;
  }

  const long long* sidx_tr() {
    if (!sidx_.p) {
      // TR with nothing recorded still needs the injection indices
      raise(KW_ERR_STATE, "time reversal needs a binary sensor mask");
    }
    return sidx_.as<long long>();
  }

  void finalize() {
    finalized_ = true;
    if (nsens_ == 0 || cartesian_) return;
    const unsigned nb = (unsigned)((nsens_ + 255) / 256);
    if (rec_ & KW_REC_P_RMS) HIPC(launch_finalize_rms(nb, stream_, p_rms_.as<double>(), nsens_, (double)Nt_));
    if (rec_ & KW_REC_U_RMS) HIPC(launch_finalize_rms(nb, stream_, u_rms_.as<double>(), nsens_, (double)Nt_));
  }

  int dev_;
  bool batch_mode_ = false;
  long long M_[3] = {1, 1, 1};   // line-FFT plan length per axis (N_, or the chirp-z plan)
  bool bs_[3] = {false, false, false};   // axis runs chirp-z (Bluestein) lines
  DevBuf chc_[3], chf_[3], chb_[3];       // chirp tables of those axes
  int nfire_ = 1;          // stacked firings (2-D), 1 otherwise
  bool plane_c2r_ = false; // 3-D step: y inverse and c2r fused per z plane
  PlaneArgs<T> plane_args_{};
  long long plane_ = 0;    // voxels of one firing's grid
  hipStream_t stream_ = nullptr;
  int nd_ = 0;
  long long N_[3] = {1, 1, 1};
  long long nvox_ = 0, NH_ = 0, NHP_ = 0;
  double dt_ = 0;
  long long Nt_ = 0, Nt_alloc_ = 0, t_ = 0;
  bool tr_ = false, stokes_ = false, finalized_ = false, cartesian_ = false, maps_f16_ = false;
  DevBuf p_, u_[3], rho_[3], rho0_, c0sq_, bona_, S1_, S2_, rho_prev_, M1_, M2_;
  DevBuf tw_[3], Dp_[3], Dm_[3], pml_[3], pmlsg_[3];
  PassCfg cfg_[3], spec_[2], c2r_;
  PassArgs<T> pa_[3];
  SpecArgs<T> sa_[2];
  C2RArgs<T> ca_;
  RecArgs<T> ra_;
  // sources
  SrcDev psrc_[3], usrc_[3];
  DevBuf pval_, uval_[3];
  long long p_ns_ = 0, p_nc_ = 1, u_ns_ = 0, u_nc_[3] = {1, 1, 1};
  bool p_dir_ = false, u_dir_ = false;
  // sensor
  long long nsens_ = 0;
  uint32_t rec_ = 0;
  DevBuf sidx_, p_ts_, u_ts_[3], p_max_, p_min_, p_rms_, p_final_, u_max_, u_rms_, u_final_, I_avg_, I_max_;
  bool dense_sensor_ = false;
  DevBuf tr_vals_;
  // adjoint
  bool het_c_ = false, nonlinear_ = false, has_p0_ = false, absorbing_ = false, no_record_ = false;
  bool axisym_ = false;
  DevBuf inv_r_;
  size_t lds_max_ = 65536;   // LDS a workgroup may allocate on this device
  long long fill_wgs_ = 512;  // workgroups that fill the device (two per CU)
  double c_ref_ = 0, tau_ = 0, y_ = 0, anp_ = 0, d_[3] = {1, 1, 1};
  std::vector<long long> psrc_idx_h_;
  std::vector<double> psrc_sig_h_, psrc_fac_h_;
  DevBuf p0_keep_;
  // post-filters
  bool dir_ = false, fr_ = false, filtered_ = false;
  double dir_size_ = 0, dir_ang_s_ = 0, fc_ = 0, bw_ = 1, dmin_ = 1;
  DevBuf dir_ang_, pfilt_;
};

template <typename T>
constexpr const char* Sim<T>::kClassNames[];

// ---------------------------------------------------------------------------
// Batch of independent simulations: every pass is one launch per line length
// across all simulations (workgroup -> simulation tables), the step index
// lives on the device, and one step can be captured in a hipGraph and
// replayed.  All simulations share precision, dimensionality and Nt.
// ---------------------------------------------------------------------------
struct BatchBase {
  virtual ~BatchBase() = default;
  virtual void run(long long n) = 0;
  virtual void sync() = 0;
  virtual SimBase* sim(int i) = 0;
  virtual int size() const = 0;
  virtual long long steps_done() const = 0;
};

template <typename T>
class Batch final : public BatchBase {
 public:
  Batch(int device, const kw_sim_desc* descs, int n, bool use_graph, bool persistent = false, int grid_cap = 0)
      : dev_(device), use_graph_(use_graph), cap_(grid_cap) {
    HIPC(hipSetDevice(dev_));
    if (persistent) {
      HIPC(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev_));
      cus_ = std::max(cus_, 1);
    }
    if (n < 1) raise(KW_ERR_INVALID, "empty batch");
    for (int i = 0; i < n; ++i) {
      if (descs[i].ndim != descs[0].ndim) raise(KW_ERR_INVALID, "a batch needs one dimensionality");
      if (descs[i].Nt != descs[0].Nt) raise(KW_ERR_INVALID, "a batch needs one Nt");
      if (descs[i].tr_data && !descs[i].axisymmetric) raise(KW_ERR_UNSUPPORTED, "time reversal is not batched");
      sims_.emplace_back(new Sim<T>(device, descs[i], true));
    }
    Nt_ = descs[0].Nt;
    nd_ = descs[0].ndim;
    HIPC(hipStreamCreate(&stream_));
    std::vector<int> t0(1, 0);
    upload(t_dev_, t0);
    build_groups();
    build_lanes();
    build_record();
  }
  ~Batch() override {
    (void)hipSetDevice(dev_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (hipEvent_t e : done_)
      if (e) (void)hipEventDestroy(e);
    if (tick_ev_) (void)hipEventDestroy(tick_ev_);
    for (hipStream_t l : lanes_)
      if (l) (void)hipStreamDestroy(l);
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  void run(long long n) override {
    HIPC(hipSetDevice(dev_));
    for (long long i = 0; i < n && t_ < Nt_; ++i) {
      if (use_graph_) {
        if (!exec_) capture();
        HIPC(hipGraphLaunch(exec_, stream_));
      } else {
        enqueue_step();
      }
      ++t_;
    }
    HIPC(hipGetLastError());
    if (t_ == Nt_ && !finalized_) {
      HIPC(hipStreamSynchronize(stream_));
      for (auto& s : sims_) {
        s->t_ = Nt_;
        s->finalize();
      }
      finalized_ = true;
    }
  }
  void sync() override {
    HIPC(hipSetDevice(dev_));
    HIPC(hipStreamSynchronize(stream_));
    for (auto& s : sims_) s->sync();
  }
  SimBase* sim(int i) override {
    if (i < 0 || i >= (int)sims_.size()) raise(KW_ERR_INVALID, "simulation index out of range");
    sync();
    sims_[i]->t_ = t_;
    return sims_[i].get();
  }
  int size() const override { return (int)sims_.size(); }
  long long steps_done() const override { return t_; }

  // Autotuner trial in batch form: `warm` untimed steps, then the mean
  // device time per step of every group over `steps` steps, reported per
  // (pass class, line length) with the lines per workgroup it ran with.
  void trial(long long warm, long long steps, std::map<std::array<int, 2>, std::pair<double, int>>& out) {
    HIPC(hipSetDevice(dev_));
    for (long long i = 0; i < warm && t_ < Nt_; ++i, ++t_) enqueue_step();
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    long long done = 0;
    for (; done < steps && t_ < Nt_; ++done, ++t_) enqueue_step(&ev);
    HIPC(hipStreamSynchronize(stream_));
    const size_t ng = groups_.size();
    std::vector<double> tot(ng, 0.0);
    for (size_t k = 0; k < ev.size(); ++k) {
      float e = 0.f;
      HIPC(hipEventElapsedTime(&e, ev[k].first, ev[k].second));
      tot[k % ng] += e;
      (void)hipEventDestroy(ev[k].first);
      (void)hipEventDestroy(ev[k].second);
    }
    for (size_t k = 0; k < ng; ++k) {
      const Group& g = *groups_[k];
      const int cls = tune_class(g.kind);
      auto& slot = out[{cls, g.n}];
      slot.first += tot[k] / (double)std::max<long long>(done, 1);
      slot.second = g.req;
    }
  }

 private:
  static int tune_class(int k) {
    switch (k) {
      case DIR_Y: return TC_DIR_Y;
      case DIR_Z: return TC_DIR_Z;
      case DIR_X: return TC_DIR_X;
      case SPEC_Z: return TC_SPEC_Z;
      case C2R: return TC_C2R;
      default: return TC_SPEC_Y;
    }
  }
  enum Kind { DIR_Y, DIR_Z, DIR_X, SPEC_Y_FWD, SPEC_Z, SPEC_Y_INV, SPEC_Y, C2R };
  struct Group {
    Kind kind;
    int n;
    int req = 1;   // lines per workgroup asked for (tuning key)
    dim3 block;
    size_t shmem;
    unsigned nwg = 0;
    bool two = false;
    DevBuf args, wg_sim, wg_first;
    std::vector<int> sims;   // simulations with this line length for this pass
    int lane = 0;            // stream lane of the concurrent step
    std::vector<int> deps;   // earlier groups on other lanes it waits for
  };

  // one group per (pass kind, line length), in step order
  void build_groups() {
    std::vector<Kind> order;
    if (nd_ == 3) order = {DIR_Y, DIR_Z, DIR_X, SPEC_Y_FWD, SPEC_Z, SPEC_Y_INV, C2R};
    else if (nd_ == 2) order = {DIR_Y, DIR_X, SPEC_Y, C2R};
    else order = {DIR_X, C2R};
    for (Kind k : order) {
      std::vector<int> lens;
      for (auto& s : sims_) {
        const int n = cfg_of(*s, k).n;
        if (std::find(lens.begin(), lens.end(), n) == lens.end()) lens.push_back(n);
      }
      for (int n : lens) {
        groups_.emplace_back(new Group);
        Group& g = *groups_.back();
        g.kind = k;
        g.n = n;
        std::vector<int> wg_sim, wg_first(sims_.size(), 0);
        std::vector<unsigned char> args;
        for (size_t i = 0; i < sims_.size(); ++i) {
          Sim<T>& s = *sims_[i];
          const PassCfg& c = cfg_of(s, k);
          if (c.n != n) continue;
          if (g.nwg > 0 && (c.block.x != g.block.x || c.shmem != g.shmem))
            raise(KW_ERR_STATE, "batched simulations disagree on the workgroup shape of a line length");
          g.block = c.block;
          g.shmem = c.shmem;
          g.req = c.req;
          g.sims.push_back((int)i);
          wg_first[i] = (int)g.nwg;
          for (unsigned w = 0; w < c.grid.x; ++w) wg_sim.push_back((int)i);
          g.nwg += c.grid.x;
          g.two = g.two || s.stokes_;
        }
        // argument records for every simulation (indexed by simulation)
        for (auto& sp : sims_) append_args(*sp, k, args);
        upload(g.args, args);
        upload(g.wg_sim, wg_sim);
        upload(g.wg_first, wg_first);
      }
    }
  }

  static const PassCfg& cfg_of(const Sim<T>& s, Kind k) {
    switch (k) {
      case DIR_Y: return s.cfg_[1];
      case DIR_Z: return s.cfg_[2];
      case DIR_X: return s.cfg_[0];
      case SPEC_Y_FWD: case SPEC_Y_INV: case SPEC_Y: return s.spec_[0];
      case SPEC_Z: return s.spec_[1];
      default: return s.c2r_;
    }
  }

  void append_args(Sim<T>& s, Kind k, std::vector<unsigned char>& out) {
    auto put = [&](const void* p, size_t n) {
      const unsigned char* b = reinterpret_cast<const unsigned char*>(p);
      out.insert(out.end(), b, b + n);
    };
    switch (k) {
      case DIR_Y: put(&s.pa_[1], sizeof(PassArgs<T>)); break;
      case DIR_Z: put(&s.pa_[2], sizeof(PassArgs<T>)); break;
      case DIR_X: put(&s.pa_[0], sizeof(PassArgs<T>)); break;
      case SPEC_Y_FWD: {
        SpecArgs<T> a = s.sa_[0];
        a.mode = 0;
        put(&a, sizeof(a));
        break;
      }
      case SPEC_Y_INV: {
        SpecArgs<T> a = s.sa_[0];
        a.mode = 1;
        a.S2 = nullptr;
        put(&a, sizeof(a));
        break;
      }
      case SPEC_Y: {
        SpecArgs<T> a = s.sa_[0];
        a.mode = 2;
        put(&a, sizeof(a));
        break;
      }
      case SPEC_Z: {
        SpecArgs<T> a = s.sa_[1];
        a.mode = 2;
        put(&a, sizeof(a));
        break;
      }
      default: put(&s.ca_, sizeof(C2RArgs<T>)); break;
    }
  }

  void build_record() {
    std::vector<int> sens_sim, sens_first(sims_.size(), 0);
    std::vector<RecArgs<T>> ra;
    for (size_t i = 0; i < sims_.size(); ++i) {
      Sim<T>& s = *sims_[i];
      ra.push_back(s.ra_);
      sens_first[i] = (int)sens_sim.size();
      if (s.nsens_ > 0 && !s.cartesian_ && s.rec_)
        for (long long j = 0; j < s.nsens_; ++j) sens_sim.push_back((int)i);
    }
    n_sens_ = (int)sens_sim.size();
    upload(rec_args_, ra);
    upload(sens_sim_, sens_sim);
    upload(sens_first_, sens_first);
  }

  // Groups touching disjoint simulations are independent, so one step is a
  // small DAG: each group depends on the latest earlier group that shares a
  // simulation with it.  Chains of groups run on up to kLanes streams (the
  // device's default count of hardware queues), the first groups of the chains
  // balanced by workgroup count, so small line lengths run beside large ones
  // instead of after them.  Lane 0 is the batch stream: it runs the step
  // counter first and the sensor gather last, after joining every lane.
  static constexpr int kLanes = 4;
  void build_lanes() {
    const int ng = (int)groups_.size();
    std::vector<int> last(sims_.size(), -1);
    std::vector<long long> load(kLanes, 0);
    std::vector<int> lane_last(kLanes, -1);
    int used = 1;
    for (int k = 0; k < ng; ++k) {
      Group& g = *groups_[k];
      std::vector<int> deps;
      for (int i : g.sims)
        if (last[i] >= 0 && std::find(deps.begin(), deps.end(), last[i]) == deps.end()) deps.push_back(last[i]);
      for (int i : g.sims) last[i] = k;
      int lane = -1;
      // continue a dependency's chain when that lane's latest group is the dependency
      for (int d : deps)
        if (lane_last[groups_[d]->lane] == d && (lane < 0 || load[groups_[d]->lane] > load[lane])) lane = groups_[d]->lane;
      if (lane < 0) {   // a new chain: the least loaded lane
        lane = 0;
        for (int l = 1; l < kLanes; ++l)
          if (load[l] < load[lane]) lane = l;
      }
      g.lane = lane;
      for (int d : deps)
        if (groups_[d]->lane != lane) g.deps.push_back(d);
      load[lane] += g.nwg;
      lane_last[lane] = k;
      used = std::max(used, lane + 1);
    }
    lanes_.assign(used, nullptr);
    lanes_[0] = stream_;
    for (int l = 1; l < used; ++l) HIPC(hipStreamCreate(&lanes_[l]));
    done_.assign(ng, nullptr);
    for (int k = 0; k < ng; ++k) HIPC(hipEventCreateWithFlags(&done_[k], hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&tick_ev_, hipEventDisableTiming));
  }

  void launch_group(const Group& g, hipStream_t st) {
    switch (g.kind) {
      case DIR_Y: case DIR_Z: case DIR_X:
        HIPC(launch_dir_batch<T>(g.n, g.kind == DIR_X, g.kind == DIR_X && g.two, g.nwg, g.block, g.shmem, st, reinterpret_cast<PassArgs<T>*>(g.args.p),
                                 reinterpret_cast<int*>(g.wg_sim.p), reinterpret_cast<int*>(g.wg_first.p), reinterpret_cast<int*>(t_dev_.p),
                                 cus_, cap_));
        break;
      case C2R:
        HIPC(launch_c2r_batch<T>(g.n, g.nwg, g.block, g.shmem, st, reinterpret_cast<C2RArgs<T>*>(g.args.p), reinterpret_cast<int*>(g.wg_sim.p),
                                 reinterpret_cast<int*>(g.wg_first.p), cus_, cap_));
        break;
      default: {
        const int mode = g.kind == SPEC_Y_FWD ? 0 : (g.kind == SPEC_Y_INV ? 1 : 2);
        HIPC(launch_spec_batch<T>(g.n, mode, g.two && mode != 1, g.nwg, g.block, g.shmem, st,
                                  reinterpret_cast<SpecArgs<T>*>(g.args.p), reinterpret_cast<int*>(g.wg_sim.p), reinterpret_cast<int*>(g.wg_first.p),
                                  cus_, cap_));
      }
    }
  }

  // One step.  With `ev` (autotuner trials) every group runs on the batch
  // stream between its own timing events; otherwise on its lane.
  void enqueue_step(std::vector<std::pair<hipEvent_t, hipEvent_t>>* ev = nullptr) {
    HIPC(launch_tick(stream_, reinterpret_cast<int*>(t_dev_.p)));
    const bool lanes = ev == nullptr && lanes_.size() > 1;
    if (lanes) HIPC(hipEventRecord(tick_ev_, stream_));
    std::vector<int> lane_last(lanes_.size(), -1);
    for (size_t k = 0; k < groups_.size(); ++k) {
      Group& g = *groups_[k];
      if (ev) {
        ev->emplace_back(nullptr, nullptr);
        HIPC(hipEventCreate(&ev->back().first));
        HIPC(hipEventCreate(&ev->back().second));
        HIPC(hipEventRecord(ev->back().first, stream_));
        launch_group(g, stream_);
        HIPC(hipEventRecord(ev->back().second, stream_));
        continue;
      }
      hipStream_t st = lanes ? lanes_[g.lane] : stream_;
      if (lanes) {
        // a lane joins the step after the counter, before its first group
        if (g.lane > 0 && lane_last[g.lane] < 0) HIPC(hipStreamWaitEvent(st, tick_ev_, 0));
        for (int d : g.deps) HIPC(hipStreamWaitEvent(st, done_[d], 0));
      }
      launch_group(g, st);
      if (lanes) {
        HIPC(hipEventRecord(done_[k], st));
        lane_last[g.lane] = (int)k;
      }
    }
    if (lanes)   // join: the gather and the next step follow every lane
      for (size_t l = 1; l < lanes_.size(); ++l)
        if (lane_last[l] >= 0) HIPC(hipStreamWaitEvent(stream_, done_[lane_last[l]], 0));
    if (n_sens_ > 0)
      HIPC(launch_record_batch<T>(stream_, reinterpret_cast<RecArgs<T>*>(rec_args_.p), reinterpret_cast<int*>(sens_sim_.p), reinterpret_cast<int*>(sens_first_.p),
                                  n_sens_, reinterpret_cast<int*>(t_dev_.p)));
  }

  void capture() {
    HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_step();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPC(hipStreamEndCapture(stream_, &graph_));
    HIPC(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }

  int dev_;
  bool use_graph_;
  int cus_ = 0;   // persistent grids: the device's compute units (0: one workgroup per tile)
  int cap_ = 0;   // persistent grids: at most this many workgroups (0: as many as fit)
  hipStream_t stream_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::vector<hipStream_t> lanes_;   // lanes_[0] == stream_
  std::vector<hipEvent_t> done_;     // per group: recorded after its launch
  hipEvent_t tick_ev_ = nullptr;
  std::vector<std::unique_ptr<Sim<T>>> sims_;
  std::vector<std::unique_ptr<Group>> groups_;
  DevBuf t_dev_, rec_args_, sens_sim_, sens_first_;
  int n_sens_ = 0, nd_ = 0;
  long long Nt_ = 0, t_ = 0;
  bool finalized_ = false;
};

// Timed trials of the lines-per-workgroup choices {default, /2, /4, /8, x2}
// for every pass class of one simulation, then, where a line length has
// alternative plan families, each family at the default and half the lines
// per workgroup; the fastest per class goes into the tuning table.  Each
// trial builds and runs its own copy of the simulation.
template <typename T>
void autotune(int dev, const kw_sim_desc& d0, int steps) {
  kw_sim_desc d = d0;
  d.tr_data = nullptr;
  const long long warm = 2;
  d.Nt = warm + steps;
  double best_ms[TC_COUNT];
  int best_lp[TC_COUNT], best_n[TC_COUNT] = {0}, best_pv[TC_COUNT];
  for (int c = 0; c < TC_COUNT; ++c) {
    best_ms[c] = INFINITY;
    best_pv[c] = -1;
  }
  struct Restore {
    ~Restore() {
      g_tune_shift = kNoTrial;
      g_tune_plan = kNoTrial;
    }
  };
  // one trial: builds the simulation with the choice in force, times it
  auto run = [&](int shift, int plan, double* ms, int* lp, int* nn, int* pv) {
    std::unique_ptr<Sim<T>> s;
    {
      Restore r;
      g_tune_shift = shift;
      g_tune_plan = plan;
      s.reset(new Sim<T>(dev, d, false));
    }
    s->trial(warm, steps, ms, lp, nn, pv);
  };
  // trials: the default, then halving (while any class still changes), then doubling
  int prev_lp[TC_COUNT] = {0}, dflt_lp[TC_COUNT] = {0};
  bool shrinking = true, has_alt = false;
  for (int shift : {0, 1, 2, 3, -1}) {
    if (shift > 1 && !shrinking) continue;
    double ms[TC_COUNT];
    int lp[TC_COUNT], nn[TC_COUNT], pv[TC_COUNT];
    run(shift, kNoTrial, ms, lp, nn, pv);
    bool changed = false;
    for (int c = 0; c < TC_COUNT; ++c) {
      if (nn[c] == 0) continue;
      if (plan_alternative(nn[c], 0)) has_alt = true;
      const int ref = shift == -1 ? dflt_lp[c] : prev_lp[c];
      if (lp[c] != ref) changed = true;
      prev_lp[c] = lp[c];
      if (shift == 0) dflt_lp[c] = lp[c];
      if (ms[c] < best_ms[c]) {
        best_ms[c] = ms[c];
        best_lp[c] = lp[c];
        best_n[c] = nn[c];
      }
    }
    if (shift > 0 && !changed) shrinking = false;   // every class already at one line per workgroup
  }
  // plan families 0 (strided), 1 (contiguous), 2 (complex-to-real) on every pass
  // whose line length has them; a class keeps the default unless another
  // family is faster (a trial reports -1 for a pass left in its default)
  for (int plan = 0; has_alt && plan <= 2; ++plan)
    for (int shift : {0, 1}) {
      double ms[TC_COUNT];
      int lp[TC_COUNT], nn[TC_COUNT], pv[TC_COUNT];
      run(shift, plan, ms, lp, nn, pv);
      for (int c = 0; c < TC_COUNT; ++c)
        if (nn[c] > 0 && pv[c] >= 0 && ms[c] < best_ms[c]) {
          best_ms[c] = ms[c];
          best_lp[c] = lp[c];
          best_n[c] = nn[c];
          best_pv[c] = pv[c];
        }
    }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int c = 0; c < TC_COUNT; ++c)
    if (best_n[c] > 0) {
      const std::array<int, 3> key{(int)sizeof(T), c, best_n[c]};
      g_tune[key] = best_lp[c];
      if (best_pv[c] >= 0) g_tune_pv[key] = best_pv[c];
      else g_tune_pv.erase(key);
    }
}

// Batch form of the autotuner: the whole batch is built and timed for each
// choice, so the table holds the shapes that are fastest when many
// simulations share each launch.
template <typename T>
void autotune_batch(int dev, const kw_sim_desc* d0, int n, int steps) {
  std::vector<kw_sim_desc> d(d0, d0 + n);
  const long long warm = 2;
  for (auto& x : d) {
    x.tr_data = nullptr;
    x.Nt = warm + steps;
  }
  std::map<std::array<int, 2>, std::pair<double, int>> best;
  std::vector<int> prev, dflt;
  // trials: the default, halving while any class still changes, then doubling
  for (int shift : {0, 1, 2, 3, -1}) {
    if (shift > 0 && shift != 1 && prev.empty()) continue;
    std::unique_ptr<Batch<T>> b;
    g_tune_shift = shift;
    try {
      b.reset(new Batch<T>(dev, d.data(), n, false));
    } catch (...) {
      g_tune_shift = kNoTrial;
      throw;
    }
    g_tune_shift = kNoTrial;
    std::map<std::array<int, 2>, std::pair<double, int>> r;
    b->trial(warm, steps, r);
    std::vector<int> lp;
    for (auto& kv : r) {
      lp.push_back(kv.second.second);
      auto it = best.find(kv.first);
      if (it == best.end() || kv.second.first < it->second.first) best[kv.first] = kv.second;
    }
    if (shift == 0) dflt = lp;
    if (shift > 0 && lp == prev) prev.clear();   // no class changes any more: skip to doubling
    else prev = lp;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& kv : best) g_tune[{(int)sizeof(T), kv.first[0], kv.first[1]}] = kv.second.second;
}

}  // namespace

struct kw_sim {
  std::unique_ptr<SimBase> impl;
};
struct kw_batch {
  std::unique_ptr<BatchBase> impl;
};

extern "C" {

int32_t kw_abi_version(void) { return KW_ABI_VERSION; }

const char* kw_last_error(void) { return g_err.c_str(); }

int32_t kw_device_count(int32_t* count) {
  KW_GUARD_BEGIN
  if (!count) raise(KW_ERR_INVALID, "null count");
  int n = 0;
  HIPC(hipGetDeviceCount(&n));
  *count = n;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_ctx_create(int32_t device, kw_ctx** out) {
  KW_GUARD_BEGIN
  if (!out) raise(KW_ERR_INVALID, "null output pointer");
  int n = 0;
  HIPC(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) raise(KW_ERR_INVALID, "device index out of range");
  HIPC(hipSetDevice(device));
  void* scratch = noalloc_scratch_address(device);
  *out = new kw_ctx{device, scratch};
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_ctx_destroy(kw_ctx* ctx) {
  delete ctx;
  return KW_OK;
}

int32_t kw_fft_length_supported(int64_t n) { return size_supported(n) ? 1 : (bluestein_plan(n) ? 2 : 0); }

int32_t kw_desc_sizes(int64_t* sim_desc_bytes, int64_t* elastic_desc_bytes) {
  if (sim_desc_bytes) *sim_desc_bytes = (int64_t)sizeof(kw_sim_desc);
  if (elastic_desc_bytes) *elastic_desc_bytes = (int64_t)sizeof(kw_elastic_desc);
  return KW_OK;
}

int32_t kw_grid_supported(int32_t ndim, const int64_t* N, int32_t precision) {
  if (ndim < 1 || ndim > 3 || N == nullptr || (precision != 4 && precision != 8)) return 0;
  long long n[3] = {1, 1, 1};
  for (int a = 0; a < ndim; ++a) {
    if (!size_supported(N[a]) && !bluestein_plan(N[a])) return 0;
    n[a] = N[a];
  }
  return lane_offsets_fit(ndim, n, precision) ? 1 : 0;
}

}  // extern "C"

namespace {

// Host storage of a stacked-firing descriptor: firings of one 2-D grid and
// medium laid one after another along a third axis.  Maps are replicated per
// firing; source and sensor indices are offset by the firing's plane; source
// signals are merged into one matrix whose column count is the longest
// firing's, shorter signals repeating their last column (the reference samples
// column min(t, ncols), first_order_steps.jl:561-715).
struct FiringStack {
  kw_sim_desc d{};
  std::vector<double> c0, rho0, bona, psig, usig[3];
  std::vector<int64_t> pidx, uidx, sidx;
};

bool same_map(const double* a, const double* b, long long n) {
  if (a == b) return true;
  if (!a || !b) return false;
  return std::memcmp(a, b, (size_t)n * sizeof(double)) == 0;
}

// Concatenate one index list of every firing (offset by the firing's plane)
// and, when `sig` is given, its signal matrix (rows x ncols, column-major).
template <typename Count, typename Index, typename Sig, typename Cols>
void merge_lists(const kw_sim_desc* ds, int n, long long plane, Count count, Index index, Sig sig, Cols ncols,
                 bool with_signal, std::vector<int64_t>& idx_out, std::vector<double>& sig_out, long long& ncols_out) {
  long long total = 0, nc = 0;
  for (int f = 0; f < n; ++f) {
    if (count(ds[f]) > 0) nc = std::max(nc, std::max<long long>(ncols(ds[f]), 1));
    total += count(ds[f]);
  }
  idx_out.clear();
  idx_out.reserve((size_t)total);
  for (int f = 0; f < n; ++f)
    for (long long j = 0; j < count(ds[f]); ++j) idx_out.push_back(index(ds[f])[j] + (int64_t)f * plane);
  ncols_out = nc;
  if (!with_signal) return;
  sig_out.assign((size_t)total * (size_t)nc, 0.0);
  long long row = 0;
  for (int f = 0; f < n; ++f) {
    const long long m = count(ds[f]), c_f = std::max<long long>(ncols(ds[f]), 1);
    const double* sg = sig(ds[f]);
    if (m > 0 && !sg) raise(KW_ERR_INVALID, "source signal missing");
    for (long long c = 0; c < nc; ++c) {
      const long long cs = std::min(c, c_f - 1);
      for (long long j = 0; j < m; ++j) sig_out[(size_t)c * total + row + j] = sg[cs * m + j];
    }
    row += m;
  }
}

void stack_firings(const kw_sim_desc* ds, int n, FiringStack& F) {
  const kw_sim_desc& a = ds[0];
  if (a.ndim != 2) raise(KW_ERR_UNSUPPORTED, "stacked firings need 2-D grids");
  const long long plane = a.N[0] * a.N[1];
  if (a.N[0] < 1 || a.N[1] < 1) raise(KW_ERR_INVALID, "grid size must be positive");
  for (int f = 0; f < n; ++f) {
    const kw_sim_desc& b = ds[f];
    const bool same = b.ndim == a.ndim && b.N[0] == a.N[0] && b.N[1] == a.N[1] && b.d[0] == a.d[0] &&
                      b.d[1] == a.d[1] && b.dt == a.dt && b.Nt == a.Nt && b.precision == a.precision &&
                      b.map_storage == a.map_storage && b.pml_size[0] == a.pml_size[0] &&
                      b.pml_size[1] == a.pml_size[1] && b.pml_alpha[0] == a.pml_alpha[0] &&
                      b.pml_alpha[1] == a.pml_alpha[1] && b.c_ref == a.c_ref && b.c0_offset == a.c0_offset &&
                      b.rho0_offset == a.rho0_offset && b.c0_scalar == a.c0_scalar &&
                      b.rho0_scalar == a.rho0_scalar && b.nonlinear == a.nonlinear &&
                      b.bona_scalar == a.bona_scalar && b.absorbing == a.absorbing &&
                      b.alpha_mode == a.alpha_mode && b.alpha_coeff == a.alpha_coeff &&
                      b.alpha_power == a.alpha_power && b.record == a.record &&
                      same_map(b.c0, a.c0, plane) && same_map(b.rho0, a.rho0, plane) &&
                      same_map(b.bona, a.bona, plane);
    if (!same) raise(KW_ERR_INVALID, "stacked firings must share the grid, time step, medium, PML and record fields");
    if (b.p0 || b.tr_data || b.directivity || b.freq_response || b.axisymmetric)
      raise(KW_ERR_UNSUPPORTED, "stacked firings take time-varying sources only (no p0, time reversal, "
                                "post-filters or axisymmetric grids)");
    if (b.n_sensor > 0 && !b.sensor_index) raise(KW_ERR_UNSUPPORTED, "stacked firings need binary sensor masks");
    if (b.n_psrc > 0 && b.p_mode != ds[0].p_mode) raise(KW_ERR_INVALID, "stacked firings need one source mode");
    if (b.n_usrc > 0 && b.u_mode != ds[0].u_mode) raise(KW_ERR_INVALID, "stacked firings need one source mode");
    for (int f2 = 0; f2 < f; ++f2)
      for (int q = 0; q < 3; ++q)
        if (b.n_usrc > 0 && ds[f2].n_usrc > 0 && (b.usrc_signal[q] == nullptr) != (ds[f2].usrc_signal[q] == nullptr))
          raise(KW_ERR_INVALID, "stacked firings need the same velocity components");
  }
  for (int f = 0; f < n; ++f)
    for (long long j = 0; j < ds[f].n_psrc + ds[f].n_usrc + ds[f].n_sensor; ++j) {
      const int64_t* ix = j < ds[f].n_psrc ? ds[f].psrc_index + j
                          : j < ds[f].n_psrc + ds[f].n_usrc ? ds[f].usrc_index + (j - ds[f].n_psrc)
                                                            : ds[f].sensor_index + (j - ds[f].n_psrc - ds[f].n_usrc);
      if (!ix || *ix < 0 || *ix >= plane) raise(KW_ERR_INVALID, "source or sensor index outside the grid");
    }
  F.d = a;
  auto rep = [&](const double* m, std::vector<double>& out) -> const double* {
    if (!m) return nullptr;
    out.resize((size_t)plane * n);
    for (int f = 0; f < n; ++f) std::memcpy(out.data() + (size_t)f * plane, m, (size_t)plane * sizeof(double));
    return out.data();
  };
  F.d.c0 = rep(a.c0, F.c0);
  F.d.rho0 = rep(a.rho0, F.rho0);
  F.d.bona = rep(a.bona, F.bona);
  long long nc = 0;
  merge_lists(
      ds, n, plane, [](const kw_sim_desc& x) { return (long long)x.n_psrc; },
      [](const kw_sim_desc& x) { return x.psrc_index; }, [](const kw_sim_desc& x) { return x.psrc_signal; },
      [](const kw_sim_desc& x) { return (long long)x.psrc_ncols; }, true, F.pidx, F.psig, nc);
  F.d.n_psrc = (int64_t)F.pidx.size();
  F.d.psrc_index = F.pidx.empty() ? nullptr : F.pidx.data();
  F.d.psrc_signal = F.pidx.empty() ? nullptr : F.psig.data();
  F.d.psrc_ncols = nc;
  for (int f = 0; f < n; ++f)
    if (ds[f].n_psrc > 0) F.d.p_mode = ds[f].p_mode;
  // velocity sources: the components the firings carry (checked equal above)
  F.d.n_usrc = 0;
  F.d.usrc_index = nullptr;
  int ref = -1;
  for (int f = n - 1; f >= 0; --f)
    if (ds[f].n_usrc > 0) ref = f;
  for (int q = 0; q < 3; ++q) {
    F.d.usrc_signal[q] = nullptr;
    F.d.usrc_ncols[q] = 0;
    if (ref < 0 || !ds[ref].usrc_signal[q]) continue;
    merge_lists(
        ds, n, plane, [](const kw_sim_desc& x) { return (long long)x.n_usrc; },
        [](const kw_sim_desc& x) { return x.usrc_index; }, [q](const kw_sim_desc& x) { return x.usrc_signal[q]; },
        [q](const kw_sim_desc& x) { return (long long)x.usrc_ncols[q]; }, true, F.uidx, F.usig[q], nc);
    F.d.usrc_signal[q] = F.usig[q].data();
    F.d.usrc_ncols[q] = nc;
  }
  if (ref >= 0) {
    F.d.u_mode = ds[ref].u_mode;
    F.d.n_usrc = (int64_t)F.uidx.size();
    F.d.usrc_index = F.uidx.data();
  }
  std::vector<double> unused;
  merge_lists(
      ds, n, plane, [](const kw_sim_desc& x) { return (long long)x.n_sensor; },
      [](const kw_sim_desc& x) { return x.sensor_index; }, [](const kw_sim_desc&) { return (const double*)nullptr; },
      [](const kw_sim_desc&) { return 1LL; }, false, F.sidx, unused, nc);
  F.d.n_sensor = (int64_t)F.sidx.size();
  F.d.sensor_index = F.sidx.empty() ? nullptr : F.sidx.data();
}

}  // namespace

extern "C" {

int32_t kw_sim_create_firings(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_firings, kw_sim** out) {
  KW_GUARD_BEGIN
  if (!ctx || !descs || !out) raise(KW_ERR_INVALID, "null argument");
  if (n_firings < 1) raise(KW_ERR_INVALID, "no firings");
  FiringStack F;
  stack_firings(descs, n_firings, F);
  std::unique_ptr<kw_sim> s(new kw_sim);
  if (F.d.precision == 4)
    s->impl.reset(new Sim<float>(ctx->device, F.d, false, n_firings));
  else if (F.d.precision == 8)
    s->impl.reset(new Sim<double>(ctx->device, F.d, false, n_firings));
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  *out = s.release();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_create(kw_ctx* ctx, const kw_sim_desc* desc, kw_sim** out) {
  KW_GUARD_BEGIN
  if (!ctx || !desc || !out) raise(KW_ERR_INVALID, "null argument");
  std::unique_ptr<kw_sim> s(new kw_sim);
  if (desc->precision == 4)
    s->impl.reset(new Sim<float>(ctx->device, *desc));
  else if (desc->precision == 8)
    s->impl.reset(new Sim<double>(ctx->device, *desc));
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  *out = s.release();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_run(kw_sim* sim, int64_t n_steps) {
  KW_GUARD_BEGIN
  if (!sim) raise(KW_ERR_INVALID, "null sim");
  if (n_steps < 0) raise(KW_ERR_INVALID, "negative step count");
  sim->impl->run(n_steps);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_synchronize(kw_sim* sim) {
  KW_GUARD_BEGIN
  if (!sim) raise(KW_ERR_INVALID, "null sim");
  sim->impl->sync();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_steps_done(const kw_sim* sim, int64_t* steps) {
  KW_GUARD_BEGIN
  if (!sim || !steps) raise(KW_ERR_INVALID, "null argument");
  *steps = sim->impl->steps_done();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_get_field(kw_sim* sim, int32_t which, double* out) {
  KW_GUARD_BEGIN
  if (!sim || !out) raise(KW_ERR_INVALID, "null argument");
  sim->impl->get_field(which, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_set_field(kw_sim* sim, int32_t which, const double* in) {
  KW_GUARD_BEGIN
  if (!sim || !in) raise(KW_ERR_INVALID, "null argument");
  sim->impl->set_field(which, in);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_sensor_shape(kw_sim* sim, uint32_t field, int64_t* rows, int64_t* cols) {
  KW_GUARD_BEGIN
  if (!sim || !rows || !cols) raise(KW_ERR_INVALID, "null argument");
  long long r = 0, c = 0;
  sim->impl->sensor_shape(field, &r, &c);
  *rows = r;
  *cols = c;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_get_sensor(kw_sim* sim, uint32_t field, double* out) {
  KW_GUARD_BEGIN
  if (!sim || !out) raise(KW_ERR_INVALID, "null argument");
  sim->impl->get_sensor(field, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_profile(kw_sim* sim, int64_t n_steps, int32_t max_classes, int32_t* n_classes, const char** names,
                       double* ms, int64_t* launches) {
  KW_GUARD_BEGIN
  if (!sim || !n_classes || !names || !ms || !launches) raise(KW_ERR_INVALID, "null argument");
  int nc = 0;
  std::vector<long long> l(max_classes > 0 ? max_classes : 1);
  sim->impl->profile(n_steps, max_classes, &nc, names, ms, l.data());
  for (int i = 0; i < nc; ++i) launches[i] = l[i];
  *n_classes = nc;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_autotune(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_sims, int32_t steps) {
  KW_GUARD_BEGIN
  if (!ctx || !descs) raise(KW_ERR_INVALID, "null argument");
  if (n_sims < 1) raise(KW_ERR_INVALID, "n_sims must be >= 1");
  if (steps < 1) steps = 1;
  HIPC(hipSetDevice(ctx->device));
  const int prec = descs[0].precision;
  if (prec != 4 && prec != 8) raise(KW_ERR_INVALID, "precision must be 4 or 8");
  for (int i = 1; i < n_sims; ++i)
    if (descs[i].precision != prec) raise(KW_ERR_INVALID, "a batch needs one precision");
  if (n_sims == 1) {
    if (prec == 4) autotune<float>(ctx->device, descs[0], steps);
    else autotune<double>(ctx->device, descs[0], steps);
  } else {
    if (prec == 4) autotune_batch<float>(ctx->device, descs, n_sims, steps);
    else autotune_batch<double>(ctx->device, descs, n_sims, steps);
  }
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_tune_export(char* buf, int64_t cap, int64_t* len) {
  KW_GUARD_BEGIN
  if (!len) raise(KW_ERR_INVALID, "null argument");
  std::ostringstream os;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& kv : g_tune) {
      os << kv.first[0] << ' ' << kv.first[1] << ' ' << kv.first[2] << ' ' << kv.second;
      auto it = g_tune_pv.find(kv.first);
      if (it != g_tune_pv.end()) os << ' ' << it->second;   // plan family, where not the default
      os << '\n';
    }
  }
  const std::string s = os.str();
  *len = (int64_t)s.size();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_tune_import(const char* text) {
  KW_GUARD_BEGIN
  if (!text) raise(KW_ERR_INVALID, "null argument");
  // one entry per line: precision bytes, class, line length, lines per
  // workgroup, and optionally the plan family (one the line length has)
  std::istringstream is(text);
  std::map<std::array<int, 3>, int> add, add_pv;
  std::string line;
  while (std::getline(is, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;   // blank line
    std::istringstream ls(line);
    long long v[6];
    int k = 0;
    while (k < 6 && ls >> v[k]) ++k;
    if (!(ls >> std::ws).eof() || (k != 4 && k != 5)) raise(KW_ERR_INVALID, "malformed tuning text");
    const long long p = v[0], c = v[1], n = v[2], l = v[3];
    if ((p != 4 && p != 8) || c < 0 || c >= TC_COUNT || n < 1 || n > (1 << 20) || !size_supported((int)n) || l < 1 ||
        l > 256)
      raise(KW_ERR_INVALID, "malformed tuning entry");
    if (k == 5 && !plan_alternative((int)n, (int)std::max(-1LL, std::min(v[4], 1000LL))))
      raise(KW_ERR_INVALID, "plan family " + std::to_string(v[4]) + " is not compiled for line length " +
                                std::to_string(n));
    const std::array<int, 3> key{(int)p, (int)c, (int)n};
    add[key] = (int)l;
    add_pv[key] = k == 5 ? (int)v[4] : -1;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& kv : add) g_tune[kv.first] = kv.second;
  for (auto& kv : add_pv)
    if (kv.second >= 0) g_tune_pv[kv.first] = kv.second;
    else g_tune_pv.erase(kv.first);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_tune_clear(void) {
  KW_GUARD_BEGIN
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  g_tune_pv.clear();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_grad_c0(kw_sim* sim, const double* p_obs, int64_t checkpoint_every, double* grad, double* loss) {
  KW_GUARD_BEGIN
  if (!sim || !p_obs || !grad || !loss) raise(KW_ERR_INVALID, "null argument");
  sim->impl->grad_c0(p_obs, checkpoint_every, grad, loss);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_sim_destroy(kw_sim* sim) {
  KW_GUARD_BEGIN
  delete sim;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_create(kw_ctx* ctx, const kw_sim_desc* descs, int32_t n_sims, int32_t use_graph, kw_batch** out) {
  KW_GUARD_BEGIN
  if (!ctx || !descs || !out) raise(KW_ERR_INVALID, "null argument");
  if (n_sims < 1) raise(KW_ERR_INVALID, "empty batch");
  for (int i = 1; i < n_sims; ++i)
    if (descs[i].precision != descs[0].precision) raise(KW_ERR_INVALID, "a batch needs one precision");
  if (use_graph < 0 || (use_graph & 0xFC) != 0) raise(KW_ERR_INVALID, "unknown batch flags");
  const bool graph = (use_graph & KW_BATCH_GRAPH) != 0, persistent = (use_graph & KW_BATCH_PERSISTENT) != 0;
  const int cap = 8 * (use_graph >> 8);   // KW_BATCH_GRID8(n): at most 8 n workgroups
  if (cap > 0 && !persistent) raise(KW_ERR_INVALID, "a grid cap needs KW_BATCH_PERSISTENT");
  std::unique_ptr<kw_batch> b(new kw_batch);
  if (descs[0].precision == 4)
    b->impl.reset(new Batch<float>(ctx->device, descs, n_sims, graph, persistent, cap));
  else if (descs[0].precision == 8)
    b->impl.reset(new Batch<double>(ctx->device, descs, n_sims, graph, persistent, cap));
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  *out = b.release();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_run(kw_batch* b, int64_t n_steps) {
  KW_GUARD_BEGIN
  if (!b) raise(KW_ERR_INVALID, "null batch");
  if (n_steps < 0) raise(KW_ERR_INVALID, "negative step count");
  b->impl->run(n_steps);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_synchronize(kw_batch* b) {
  KW_GUARD_BEGIN
  if (!b) raise(KW_ERR_INVALID, "null batch");
  b->impl->sync();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_sensor_shape(kw_batch* b, int32_t sim, uint32_t field, int64_t* rows, int64_t* cols) {
  KW_GUARD_BEGIN
  if (!b || !rows || !cols) raise(KW_ERR_INVALID, "null argument");
  long long r = 0, c = 0;
  b->impl->sim(sim)->sensor_shape(field, &r, &c);
  *rows = r;
  *cols = c;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_get_sensor(kw_batch* b, int32_t sim, uint32_t field, double* out) {
  KW_GUARD_BEGIN
  if (!b || !out) raise(KW_ERR_INVALID, "null argument");
  b->impl->sim(sim)->get_sensor(field, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_get_field(kw_batch* b, int32_t sim, int32_t which, double* out) {
  KW_GUARD_BEGIN
  if (!b || !out) raise(KW_ERR_INVALID, "null argument");
  b->impl->sim(sim)->get_field(which, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_batch_destroy(kw_batch* b) {
  KW_GUARD_BEGIN
  delete b;
  return KW_OK;
  KW_GUARD_END
}

}  // extern "C"
