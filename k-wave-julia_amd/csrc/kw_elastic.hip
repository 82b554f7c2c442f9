// C ABI of the elastic stress-velocity PSTD solver (pstd_elastic_2d /
// pstd_elastic_3d, KWave.jl/src/solver/elastic.jl:133-588): host tables,
// device fields and the pass sequence of one time step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "kw_host.hpp"
#include "kw_launch.hpp"

using namespace kw;
using namespace kwh;

namespace {

template <typename T>
class ElasticSim {
 public:
  ElasticSim(int device, const kw_elastic_desc& d) : dev_(device) { build(d); }
  ~ElasticSim() {
    if (stream_) {
      (void)hipSetDevice(dev_);
      (void)hipStreamSynchronize(stream_);
      (void)hipStreamDestroy(stream_);
    }
  }

  void run(long long n) {
    HIPC(hipSetDevice(dev_));
    for (long long i = 0; i < n && t_ < Nt_; ++i) step();
    HIPC(hipGetLastError());
    if (t_ == Nt_ && !finalized_) finalize();
  }
  void sync() {
    HIPC(hipSetDevice(dev_));
    HIPC(hipStreamSynchronize(stream_));
  }
  long long steps_done() const { return t_; }

  void get_field(int which, double* out) {
    HIPC(hipSetDevice(dev_));
    sync();
    const T* src = nullptr;
    if (which >= KW_EL_VX && which <= KW_EL_VZ && which - KW_EL_VX < nd_)
      src = v_[which - KW_EL_VX].template as<T>();
    else if (which >= KW_EL_SXX && which <= KW_EL_SZZ && which - KW_EL_SXX < nd_)
      src = sn_[which - KW_EL_SXX].template as<T>();
    else if (which == KW_EL_SXY)
      src = ss_[0].template as<T>();
    else if ((which == KW_EL_SXZ || which == KW_EL_SYZ) && nd_ == 3)
      src = ss_[which - KW_EL_SXY].template as<T>();
    else
      raise(KW_ERR_INVALID, "invalid elastic field selector");
    copy_out(src, nvox_, out);
  }

  // p = -tr(s)/ndim over the whole grid (elastic.jl:350, 573)
  void get_pressure(double* out) {
    HIPC(hipSetDevice(dev_));
    sync();
    std::vector<double> s(nvox_);
    for (long long i = 0; i < nvox_; ++i) out[i] = 0.0;
    for (int a = 0; a < nd_; ++a) {
      copy_out(sn_[a].template as<T>(), nvox_, s.data());
      for (long long i = 0; i < nvox_; ++i) out[i] += s[i];
    }
    for (long long i = 0; i < nvox_; ++i) out[i] = -out[i] / (double)nd_;
  }

  // per-kernel-class mean device time over n steps (HIP events on the stream)
  void profile(long long n, int maxc, int* nc, const char** names, double* ms, long long* launches) {
    HIPC(hipSetDevice(dev_));
    std::vector<ProfRec> ev;
    ev.reserve(64);
    prof_ = &ev;
    for (long long i = 0; i < n && t_ < Nt_; ++i) step();
    prof_ = nullptr;
    HIPC(hipStreamSynchronize(stream_));
    double tot[kNumClasses] = {};
    long long cnt[kNumClasses] = {};
    for (auto& r : ev) {
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

  void sensor_shape(uint32_t f, long long* r, long long* c) {
    if (!(rec_ & f)) raise(KW_ERR_INVALID, "field not recorded");
    *r = nsens_;
    *c = (f == KW_REC_P || f == KW_REC_UX || f == KW_REC_UY || f == KW_REC_UZ) ? Nt_ : 1;
  }

  void get_sensor(uint32_t f, double* out) {
    HIPC(hipSetDevice(dev_));
    sync();
    long long rows, cols;
    sensor_shape(f, &rows, &cols);
    if (cartesian_) {
      // Cartesian masks are preallocated and never filled (elastic.jl:210-215)
      double init = 0.0;
      if (f == KW_REC_P_MAX || f == KW_REC_U_MAX || f == KW_REC_I_MAX) init = -INFINITY;
      if (f == KW_REC_P_MIN) init = INFINITY;
      for (long long i = 0; i < rows * cols; ++i) out[i] = init;
      return;
    }
    switch (f) {
      case KW_REC_P:
        copy_out(p_ts_.as<T>(), rows * cols, out);
        break;
      case KW_REC_UX:
      case KW_REC_UY:
      case KW_REC_UZ: {
        const int d = f == KW_REC_UX ? 0 : (f == KW_REC_UY ? 1 : 2);
        if (d >= nd_)
          for (long long i = 0; i < rows * cols; ++i) out[i] = 0.0;
        else
          copy_out(u_ts_[d].as<T>(), rows * cols, out);
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

 private:
  void copy_out(const T* dsrc, long long n, double* out) {
    std::vector<T> h(n);
    if (n) HIPC(hipMemcpy(h.data(), dsrc, n * sizeof(T), hipMemcpyDeviceToHost));
    for (long long i = 0; i < n; ++i) out[i] = (double)h[i];
  }
  void zeros(DevBuf& b) {
    b.alloc(nvox_ * sizeof(T));
    HIPC(hipMemset(b.p, 0, b.bytes));
  }

  // ---- construction -------------------------------------------------------
  void build(const kw_elastic_desc& d) {
    HIPC(hipSetDevice(dev_));
    nd_ = d.ndim;
    if (nd_ != 2 && nd_ != 3) raise(KW_ERR_INVALID, "pstd_elastic needs a 2-D or 3-D grid");
    for (int a = 0; a < 3; ++a) N_[a] = a < nd_ ? d.N[a] : 1;
    for (int a = 0; a < nd_; ++a) {
      if (N_[a] < 1) raise(KW_ERR_INVALID, "grid size must be positive");
      if (!size_supported(N_[a]))
        raise(KW_ERR_UNSUPPORTED, "axis length " + std::to_string(N_[a]) + " has no line-FFT plan");
      if (d.pml_size[a] < 0 || d.pml_size[a] > N_[a]) raise(KW_ERR_INVALID, "pml_size out of range");
    }
    if (d.dt == 0.0 || d.Nt == 0)
      raise(KW_ERR_INVALID, "Time array not set. Call make_time!(kgrid, sound_speed) before running.");
    {
      int v = 0;
      HIPC(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_));
      lds_max_ = v > 0 ? (size_t)v : 65536;
      int cus = 0;
      HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_));
      fill_wgs_ = 2LL * std::max(cus, 1);   // two workgroups per CU
    }
    check_lane_offsets(nd_, N_, (int)sizeof(T));
    nvox_ = N_[0] * N_[1] * N_[2];
    dt_ = d.dt;
    Nt_ = d.Nt;
    HIPC(hipStreamCreate(&stream_));

    // fields: velocities, normal stresses (-p0 at t = 0), shear stresses
    for (int a = 0; a < nd_; ++a) zeros(v_[a]);
    for (int a = 0; a < nd_; ++a) {
      if (d.p0) {
        std::vector<double> s(nvox_);
        for (long long i = 0; i < nvox_; ++i) s[i] = -d.p0[i];
        upload(sn_[a], to_t<T>(s));
      } else {
        zeros(sn_[a]);
      }
    }
    npairs_ = nd_ == 2 ? 1 : 3;
    for (int k = 0; k < npairs_; ++k) zeros(ss_[k]);
    zeros(scr_[0]);
    if (nd_ == 3) zeros(scr_[1]);

    // medium: Lame parameters lambda = rho (cp^2 - 2 cs^2), mu = rho cs^2 (elastic.jl:186-190)
    const bool het = d.cp || d.cs || d.rho0;
    auto at = [&](const double* m, double s, long long i) { return m ? m[i] : s; };
    if (het) {
      std::vector<double> lam(nvox_), mu(nvox_);
      for (long long i = 0; i < nvox_; ++i) {
        const double cp = at(d.cp, d.cp_scalar, i), cs = at(d.cs, d.cs_scalar, i), r = at(d.rho0, d.rho0_scalar, i);
        lam[i] = r * (cp * cp - 2.0 * cs * cs);
        mu[i] = r * cs * cs;
      }
      upload(lam_, to_t<T>(lam));
      upload(mu_, to_t<T>(mu));
    }
    lam_s_ = d.rho0_scalar * (d.cp_scalar * d.cp_scalar - 2.0 * d.cs_scalar * d.cs_scalar);
    mu_s_ = d.rho0_scalar * d.cs_scalar * d.cs_scalar;
    if (d.rho0) upload(rho_, to_t<T>(std::vector<double>(d.rho0, d.rho0 + nvox_)));
    rho_s_ = d.rho0_scalar;
    // the reference's array branch carries no viscous terms (elastic.jl:323-334)
    kp_ = het ? 1.0 : 1.0 + d.chi_p;
    ks_ = het ? 1.0 : 1.0 + d.chi_s;

    for (int a = 0; a < nd_; ++a) {
      AxisTables A = axis_tables(N_[a], d.d[a], d.pml_size[a], d.pml_alpha[a], d.dt);
      upload(tw_[a], to_cx<T>(A.twr, A.twi));
      upload(Dp_[a], to_cx<T>(A.dpr, A.dpi));
      upload(Dm_[a], to_cx<T>(A.dmr, A.dmi));
      upload(pml_[a], to_t<T>(A.pml));
      upload(pmlsg_[a], to_t<T>(A.pml_sg));
    }
    cfg_[0] = line_cfg(N_[0], true, N_[1] * N_[2], 0);
    cfg_[1] = line_cfg(N_[1], false, N_[0], N_[2]);
    if (nd_ == 3) cfg_[2] = line_cfg(N_[2], false, N_[0], N_[1]);
    for (int a = 0; a < nd_; ++a) base_[a] = base_args(a);

    build_sources(d);
    build_sensor(d);
    HIPC(hipStreamSynchronize(stream_));
  }

  PassCfg line_cfg(long long n, bool contig, long long lines, long long outer) {
    return line_pass_cfg(n, contig, lines, outer, sizeof(Cx<T>), lds_max_, fill_wgs_);
  }

  ElArgs<T> base_args(int axis) {
    ElArgs<T> A{};
    const PassCfg& c = cfg_[axis];
    A.tw = tw_[axis].template as<Cx<T>>();
    A.Nx = (int)N_[0];
    A.Ny = (int)N_[1];
    A.n_lines = (int)(N_[1] * N_[2]);
    A.x_blocks = (int)(((N_[0] + 1) / 2 + c.lpw - 1) / c.lpw);
    if (axis == 0) {
      A.Sa = 1;
      A.So = N_[0];
    } else if (axis == 1) {
      A.Sa = N_[0];
      A.So = N_[0] * N_[1];
    } else {
      A.Sa = N_[0] * N_[1];
      A.So = N_[0];
    }
    A.axis = axis;
    A.ndim = nd_;
    A.dt = (T)dt_;
    for (int b = 0; b < 3; ++b) A.pml[b] = b < nd_ ? pml_[b].template as<T>() : nullptr;
    A.lam = lam_.template as<T>();
    A.lam_s = (T)lam_s_;
    A.mu = mu_.template as<T>();
    A.mu_s = (T)mu_s_;
    A.rho = rho_.template as<T>();
    A.rho_s = (T)rho_s_;
    A.pml_sg = pmlsg_[axis].template as<T>();
    A.kp = (T)kp_;
    A.ks = (T)ks_;
    return A;
  }

  void build_sources(const kw_elastic_desc& d) {
    auto signals = [&](long long ns, const int64_t* index, const double* sig, long long nc, DevBuf& out) {
      if (nc < 1) raise(KW_ERR_INVALID, "source signal needs at least one column");
      std::vector<T> vals(ns * nc);
      for (long long j = 0; j < ns; ++j)
        for (long long k = 0; k < nc; ++k) vals[k * ns + j] = (T)sig[j + ns * k];
      upload(out, vals);
      (void)index;
    };
    auto indices = [&](long long ns, const int64_t* index, DevBuf& out) {
      std::vector<long long> idx(index, index + ns);
      for (long long v : idx)
        if (v < 0 || v >= nvox_) raise(KW_ERR_INVALID, "source index out of range");
      upload(out, idx);
    };
    if (d.n_usrc > 0 && d.usrc_index) {
      u_ns_ = d.n_usrc;
      indices(u_ns_, d.usrc_index, uidx_);
      for (int a = 0; a < nd_; ++a)
        if (d.usrc_signal[a]) {
          signals(u_ns_, d.usrc_index, d.usrc_signal[a], d.usrc_ncols[a], uval_[a]);
          u_nc_[a] = d.usrc_ncols[a];
        }
      u_dir_ = d.u_mode == KW_DIRICHLET;
    }
    if (d.n_ssrc > 0 && d.ssrc_index) {
      s_ns_ = d.n_ssrc;
      indices(s_ns_, d.ssrc_index, sidx_src_);
      for (int k = 0; k < 6; ++k) {
        const bool normal = k < 3;
        if (!d.ssrc_signal[k]) continue;
        if (normal && k >= nd_) continue;
        if (!normal && nd_ == 2 && k != 3) continue;   // 2-D has sxy only
        signals(s_ns_, d.ssrc_index, d.ssrc_signal[k], d.ssrc_ncols[k], sval_[k]);
        s_nc_[k] = d.ssrc_ncols[k];
      }
      s_dir_ = d.s_mode == KW_DIRICHLET;
    }
  }

  void build_sensor(const kw_elastic_desc& d) {
    nsens_ = d.n_sensor;
    rec_ = d.record;
    cartesian_ = d.n_sensor > 0 && d.sensor_index == nullptr;
    if (nsens_ == 0 || cartesian_) return;
    std::vector<long long> idx(d.sensor_index, d.sensor_index + nsens_);
    for (long long v : idx)
      if (v < 0 || v >= nvox_) raise(KW_ERR_INVALID, "sensor index out of range");
    upload(sens_idx_, idx);
    zeros(p_);
    auto fill_t = [&](DevBuf& b, long long n, double v) { upload(b, std::vector<T>(n, (T)v)); };
    auto fill_d = [&](DevBuf& b, long long n, double v) { upload(b, std::vector<double>(n, v)); };
    if (rec_ & KW_REC_P) fill_t(p_ts_, nsens_ * Nt_, 0.0);
    for (int a = 0; a < nd_; ++a) {
      const uint32_t bit = a == 0 ? KW_REC_UX : (a == 1 ? KW_REC_UY : KW_REC_UZ);
      if (rec_ & bit) fill_t(u_ts_[a], nsens_ * Nt_, 0.0);
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
    ra_ = RecArgs<T>{};
    ra_.p = p_.as<T>();
    for (int a = 0; a < 3; ++a) ra_.u[a] = a < nd_ ? v_[a].template as<T>() : nullptr;
    ra_.ndim = nd_;
    ra_.idx = sens_idx_.as<long long>();
    ra_.n = (int)nsens_;
    ra_.Nt = (int)Nt_;
    ra_.rec = rec_;
    ra_.p_ts = p_ts_.as<T>();
    for (int a = 0; a < 3; ++a) ra_.u_ts[a] = u_ts_[a].template as<T>();
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

  // ---- one time step ------------------------------------------------------
  T* s_of(int a, int b) {
    if (a == b) return sn_[a].template as<T>();
    const int lo = std::min(a, b), hi = std::max(a, b);
    const int k = lo == 0 ? hi - 1 : 2;   // (0,1) -> 0, (0,2) -> 1, (1,2) -> 2
    return ss_[k].template as<T>();
  }
  // kernel classes: mode x line axis, then the small kernels
  static constexpr int kNumClasses = 14;
  static constexpr const char* kClassNames[kNumClasses] = {
      "k_el_deriv_x", "k_el_deriv_y", "k_el_deriv_z", "k_el_vel_x",  "k_el_vel_y",   "k_el_vel_z",   "k_el_normal_x",
      "k_el_normal_y", "k_el_normal_z", "k_el_shear_x", "k_el_shear_y", "k_el_shear_z", "k_el_inject", "k_el_record"};
  enum { CLS_INJECT = 12, CLS_RECORD = 13 };
  struct ProfRec {
    int cls;
    hipEvent_t a, b;
  };
  template <typename F>
  void timed(int cls, F&& f) {
    if (!prof_) {
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
    prof_->push_back(r);
  }
  void launch(int axis, const ElArgs<T>& A) {
    const PassCfg& c = cfg_[axis];
    timed(3 * A.mode + axis, [&] { HIPC(launch_el<T>(c.n, axis == 0, c.grid, c.block, c.shmem, stream_, A)); });
  }
  // out (+)= D_axis f
  void deriv(int axis, const T* f, const Cx<T>* D, T* out, bool accumulate) {
    ElArgs<T> A = base_[axis];
    A.mode = EL_DERIV;
    A.f = f;
    A.D = D;
    A.out = out;
    A.accumulate = accumulate ? 1 : 0;
    launch(axis, A);
  }

  void step() {
    const int t = (int)t_ + 1;
    T* s0 = scr_[0].template as<T>();
    T* s1 = scr_[1].template as<T>();
    // momentum (elastic.jl:228-246, 470-500): v_a from sum_b D_b+ s_ab
    for (int a = 0; a < nd_; ++a) {
      bool first = true;
      for (int b = 0; b < nd_; ++b) {
        if (b == a) continue;
        deriv(b, s_of(a, b), Dp_[b].template as<Cx<T>>(), s0, !first);
        first = false;
      }
      ElArgs<T> A = base_[a];
      A.mode = EL_VEL;
      A.f = s_of(a, a);
      A.D = Dp_[a].template as<Cx<T>>();
      A.acc = s0;
      A.v = v_[a].template as<T>();
      launch(a, A);
    }
    // velocity source (elastic.jl:248-275)
    for (int a = 0; a < nd_; ++a) {
      if (!uval_[a].p) continue;
      const long long col = std::min<long long>(t, u_nc_[a]) - 1;
      timed(CLS_INJECT, [&] {
        HIPC(launch_el_inject<T>(stream_, v_[a].template as<T>(), uidx_.as<long long>(),
                                 uval_[a].template as<T>() + col * u_ns_, (int)u_ns_, u_dir_ ? 1 : 0));
      });
    }
    // normal stresses (elastic.jl:288-316, 540-547): D_a- v_a
    deriv(0, v_[0].template as<T>(), Dm_[0].template as<Cx<T>>(), s0, false);
    if (nd_ == 3) deriv(1, v_[1].template as<T>(), Dm_[1].template as<Cx<T>>(), s1, false);
    {
      const int last = nd_ - 1;
      ElArgs<T> A = base_[last];
      A.mode = EL_NORMAL;
      A.f = v_[last].template as<T>();
      A.D = Dm_[last].template as<Cx<T>>();
      A.dn[0] = s0;
      A.dn[1] = s1;
      for (int k = 0; k < nd_; ++k) A.s[k] = sn_[k].template as<T>();
      launch(last, A);
    }
    // shear stresses s_ab = P_a P_b (s_ab + dt mu (D_b- v_a + D_a- v_b))
    for (int a = 0; a < nd_; ++a)
      for (int b = a + 1; b < nd_; ++b) {
        deriv(b, v_[a].template as<T>(), Dm_[b].template as<Cx<T>>(), s0, false);
        ElArgs<T> A = base_[a];
        A.mode = EL_SHEAR;
        A.f = v_[b].template as<T>();
        A.D = Dm_[a].template as<Cx<T>>();
        A.acc = s0;
        A.axis2 = b;
        A.s[0] = s_of(a, b);
        launch(a, A);
      }
    // stress source (elastic.jl:318-345, 552-570): sxx syy szz sxy sxz syz
    for (int k = 0; k < 6; ++k) {
      if (!sval_[k].p) continue;
      T* f = k < 3 ? sn_[k].template as<T>() : ss_[k - 3].template as<T>();
      const long long col = std::min<long long>(t, s_nc_[k]) - 1;
      timed(CLS_INJECT, [&] {
        HIPC(launch_el_inject<T>(stream_, f, sidx_src_.as<long long>(), sval_[k].template as<T>() + col * s_ns_,
                                 (int)s_ns_, s_dir_ ? 1 : 0));
      });
    }
    // record p = -tr(s)/ndim and the velocities at the sensor points
    if (nsens_ > 0 && !cartesian_) {
      ra_.t = t;
      timed(CLS_RECORD, [&] {
        HIPC(launch_el_sensor_p<T>(stream_, p_.as<T>(), sn_[0].template as<T>(), sn_[1].template as<T>(),
                                   nd_ == 3 ? sn_[2].template as<T>() : nullptr, sens_idx_.as<long long>(),
                                   (int)nsens_, nd_));
        HIPC(launch_record<T>((unsigned)((nsens_ + 255) / 256), stream_, ra_));
      });
    }
    t_ = t;
  }

  void finalize() {
    finalized_ = true;
    if (nsens_ == 0 || cartesian_) return;
    const unsigned nb = (unsigned)((nsens_ + 255) / 256);
    if (rec_ & KW_REC_P_RMS) HIPC(launch_finalize_rms(nb, stream_, p_rms_.as<double>(), nsens_, (double)Nt_));
    if (rec_ & KW_REC_U_RMS) HIPC(launch_finalize_rms(nb, stream_, u_rms_.as<double>(), nsens_, (double)Nt_));
  }

  int dev_;
  hipStream_t stream_ = nullptr;
  int nd_ = 0, npairs_ = 1;
  long long N_[3] = {1, 1, 1};
  long long nvox_ = 0, Nt_ = 0, t_ = 0;
  double dt_ = 0;
  bool finalized_ = false, cartesian_ = false;
  size_t lds_max_ = 65536;
  long long fill_wgs_ = 512;
  DevBuf v_[3], sn_[3], ss_[3], scr_[2], p_;
  DevBuf lam_, mu_, rho_;
  double lam_s_ = 0, mu_s_ = 0, rho_s_ = 1, kp_ = 1, ks_ = 1;
  DevBuf tw_[3], Dp_[3], Dm_[3], pml_[3], pmlsg_[3];
  PassCfg cfg_[3];
  ElArgs<T> base_[3];
  // sources
  DevBuf uidx_, uval_[3], sidx_src_, sval_[6];
  long long u_ns_ = 0, s_ns_ = 0, u_nc_[3] = {1, 1, 1}, s_nc_[6] = {1, 1, 1, 1, 1, 1};
  bool u_dir_ = false, s_dir_ = false;
  // sensor
  long long nsens_ = 0;
  uint32_t rec_ = 0;
  DevBuf sens_idx_, p_ts_, u_ts_[3], p_max_, p_min_, p_rms_, p_final_, u_max_, u_rms_, u_final_, I_avg_, I_max_;
  RecArgs<T> ra_{};
  std::vector<ProfRec>* prof_ = nullptr;
};

template <typename T>
constexpr const char* ElasticSim<T>::kClassNames[];

struct ElBase {
  virtual ~ElBase() = default;
  virtual void run(long long n) = 0;
  virtual void sync() = 0;
  virtual long long steps_done() const = 0;
  virtual void get_field(int which, double* out) = 0;
  virtual void get_pressure(double* out) = 0;
  virtual void sensor_shape(uint32_t f, long long* r, long long* c) = 0;
  virtual void get_sensor(uint32_t f, double* out) = 0;
  virtual void profile(long long n, int maxc, int* nc, const char** names, double* ms, long long* launches) = 0;
};
template <typename T>
struct ElImpl final : ElBase {
  explicit ElImpl(int dev, const kw_elastic_desc& d) : s(dev, d) {}
  void run(long long n) override { s.run(n); }
  void sync() override { s.sync(); }
  long long steps_done() const override { return s.steps_done(); }
  void get_field(int which, double* out) override { s.get_field(which, out); }
  void get_pressure(double* out) override { s.get_pressure(out); }
  void sensor_shape(uint32_t f, long long* r, long long* c) override { s.sensor_shape(f, r, c); }
  void get_sensor(uint32_t f, double* out) override { s.get_sensor(f, out); }
  void profile(long long n, int maxc, int* nc, const char** names, double* ms, long long* launches) override {
    s.profile(n, maxc, nc, names, ms, launches);
  }
  ElasticSim<T> s;
};

}  // namespace

struct kw_elastic {
  std::unique_ptr<ElBase> impl;
};

extern "C" {

int32_t kw_elastic_create(kw_ctx* ctx, const kw_elastic_desc* desc, kw_elastic** out) {
  KW_GUARD_BEGIN
  if (!ctx || !desc || !out) raise(KW_ERR_INVALID, "null argument");
  std::unique_ptr<kw_elastic> s(new kw_elastic);
  if (desc->precision == 4)
    s->impl.reset(new ElImpl<float>(ctx->device, *desc));
  else if (desc->precision == 8)
    s->impl.reset(new ElImpl<double>(ctx->device, *desc));
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  *out = s.release();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_run(kw_elastic* sim, int64_t n_steps) {
  KW_GUARD_BEGIN
  if (!sim) raise(KW_ERR_INVALID, "null sim");
  if (n_steps < 0) raise(KW_ERR_INVALID, "negative step count");
  sim->impl->run(n_steps);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_synchronize(kw_elastic* sim) {
  KW_GUARD_BEGIN
  if (!sim) raise(KW_ERR_INVALID, "null sim");
  sim->impl->sync();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_steps_done(const kw_elastic* sim, int64_t* steps) {
  KW_GUARD_BEGIN
  if (!sim || !steps) raise(KW_ERR_INVALID, "null argument");
  *steps = sim->impl->steps_done();
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_get_field(kw_elastic* sim, int32_t which, double* out) {
  KW_GUARD_BEGIN
  if (!sim || !out) raise(KW_ERR_INVALID, "null argument");
  sim->impl->get_field(which, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_get_pressure(kw_elastic* sim, double* out) {
  KW_GUARD_BEGIN
  if (!sim || !out) raise(KW_ERR_INVALID, "null argument");
  sim->impl->get_pressure(out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_sensor_shape(kw_elastic* sim, uint32_t field, int64_t* rows, int64_t* cols) {
  KW_GUARD_BEGIN
  if (!sim || !rows || !cols) raise(KW_ERR_INVALID, "null argument");
  long long r = 0, c = 0;
  sim->impl->sensor_shape(field, &r, &c);
  *rows = r;
  *cols = c;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_get_sensor(kw_elastic* sim, uint32_t field, double* out) {
  KW_GUARD_BEGIN
  if (!sim || !out) raise(KW_ERR_INVALID, "null argument");
  sim->impl->get_sensor(field, out);
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_profile(kw_elastic* sim, int64_t n_steps, int32_t max_classes, int32_t* n_classes,
                           const char** names, double* ms, int64_t* launches) {
  KW_GUARD_BEGIN
  if (!sim || !n_classes || !names || !ms || !launches) raise(KW_ERR_INVALID, "null argument");
  if (n_steps < 0) raise(KW_ERR_INVALID, "negative step count");
  int nc = 0;
  std::vector<long long> l(std::max(0, (int)max_classes));
  sim->impl->profile(n_steps, max_classes, &nc, names, ms, l.data());
  for (int i = 0; i < nc; ++i) launches[i] = l[i];
  *n_classes = nc;
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_elastic_destroy(kw_elastic* sim) {
  delete sim;
  return KW_OK;
}

}  // extern "C"
