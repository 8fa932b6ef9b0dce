// pybind_module.cpp -- the drop-in `rils_rols_cpp` extension module.
//
// Same module name, class and method signatures as the reference's
// PYBIND11_MODULE (rils_rols_cpp/rils_rols_cpp.cpp:996-1007).  Differences:
//   * input-size mismatches raise ValueError instead of printing and calling
//     exit(1) (rils_rols_cpp.cpp:679-682, 704-707, 722-725, 733-736);
//   * the GIL is released while fitting;
//   * extra methods (moves, stats, precision, device) and test hooks prefixed
//     with an underscore.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>

#include "search.h"
#include "serve.h"
#include "tree.h"

namespace py = pybind11;
using rrh::Flat;
using rrh::RilsRols;

namespace {

py::dict stats_dict(const rrh::Stats& s) {
  py::dict d;
  d["sweeps"] = s.sweeps; d["batched_cands"] = s.batched_cands; d["raw_batches"] = s.raw_batches;
  d["raw_cands"] = s.raw_cands; d["single_calls"] = s.single_calls; d["rescored_fp64"] = s.rescored_fp64;
  d["host_gen_s"] = s.host_gen_s; d["host_prep_s"] = s.host_prep_s; d["gpu_s"] = s.gpu_s;
  d["host_main_s"] = s.host_main_s; d["setup_s"] = s.setup_s;
  d["kernel_ms"] = s.kernel_ms; d["replay_s"] = s.replay_s;
  return d;
}

py::list moves_list(const std::vector<rrh::Move>& mv) {
  py::list out;
  for (const rrh::Move& m : mv) out.append(py::make_tuple(m.kind, m.fit_calls, m.e0, m.e1, m.size, m.model));
  return out;
}

// merge hook for tests: batches given as (cand_start, prog_start, code, consts, dataset) tuples
py::tuple merge_hook(const std::vector<std::tuple<std::vector<int32_t>, std::vector<int32_t>, std::vector<uint32_t>,
                                                  std::vector<double>, int32_t>>& parts) {
  std::vector<rr_batch> bs(parts.size());
  std::vector<rrh::BatchView> views;
  for (size_t i = 0; i < parts.size(); ++i) {
    const auto& t = parts[i];
    rr_batch& b = bs[i];
    b.n_cands = (int32_t)std::get<0>(t).size() - 1;
    b.cand_start = std::get<0>(t).data();
    b.prog_start = std::get<1>(t).data();
    b.code = std::get<2>(t).data();
    b.consts = std::get<3>(t).data();
    b.n_consts = (int32_t)std::get<3>(t).size();
    views.push_back({&b, std::get<4>(t)});
  }
  std::vector<int32_t> cs, ps, hs;
  std::vector<uint32_t> code;
  std::vector<double> consts;
  rrh::merge_batches(views, cs, ps, code, consts, hs);
  return py::make_tuple(cs, ps, code, consts, hs);
}

py::array_t<double> as_f64(py::array_t<double, py::array::c_style | py::array::forcecast> a) { return a; }

std::vector<std::string> candidates_hook(const std::string& prefix, int nvars, bool classification, bool local_search) {
  RilsRols rr(classification, 1, 1, 0.001, 50, 1.0, false, 0);
  std::vector<int> rel(nvars);
  for (int i = 0; i < nvars; ++i) rel[i] = i;
  rr.setup_nodes(rel);
  std::vector<Flat> c = rr.all_candidates(rrh::parse_prefix(prefix), local_search);
  std::vector<std::string> out;
  out.reserve(c.size());
  for (const Flat& f : c) out.push_back(rrh::to_prefix(f));
  return out;
}

// relevant_features of a fresh search over X (n x d row-major) and y
std::vector<int> features_hook(py::array_t<double, py::array::c_style | py::array::forcecast> X,
                               py::array_t<double, py::array::c_style | py::array::forcecast> y) {
  py::buffer_info bx = X.request(), by = y.request();
  if (bx.ndim != 2 || by.ndim != 1 || by.shape[0] != bx.shape[0]) throw std::invalid_argument("X must be n x d, y n");
  const int64_t n = bx.shape[0], d = bx.shape[1];
  const double* xp = (const double*)bx.ptr;
  std::vector<std::vector<double>> Xf(d, std::vector<double>(n));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < d; ++j) Xf[j][i] = xp[i * d + j];
  std::vector<double> yv((const double*)by.ptr, (const double*)by.ptr + n);
  RilsRols rr(false, 1, 1, 0.001, 50, 1.0, false, 0);
  return rr.relevant_features(Xf, yv);
}

py::object rewrite_hook(const std::string& prefix, const std::string& op) {
  Flat f = rrh::parse_prefix(prefix);
  rrh::Pool pool;
  if (op == "to_string") return py::str(rrh::to_string(f));
  if (op == "size") return py::int_((int)f.size());
  if (op == "candidate_rounds") { rrh::rewrite_rounds(f, true, pool); return py::str(rrh::to_prefix(f)); }
  if (op == "accept_rounds") { rrh::rewrite_rounds(f, false, pool); return py::str(rrh::to_prefix(f)); }
  const int32_t root = pool.load(f);
  if (op == "expand") pool.expand(root);
  else if (op == "simplify") pool.simplify(root);
  else if (op == "normalize") pool.normalize_factor_constants(root, false);
  else throw std::invalid_argument("unknown rewrite op " + op);
  Flat out;
  pool.store(root, out);
  return py::str(rrh::to_prefix(out));
}

std::vector<std::string> factors_hook(const std::string& prefix) {
  rrh::Pool pool;
  std::vector<Flat> fs;
  rrh::ols_factors(rrh::parse_prefix(prefix), fs, pool);
  std::vector<std::string> out;
  for (const Flat& f : fs) out.push_back(rrh::to_prefix(f));
  return out;
}

py::tuple compile_hook(const std::string& prefix) {
  std::vector<uint32_t> code;
  std::vector<double> consts;
  const int depth = rrh::compile(rrh::parse_prefix(prefix), code, consts);
  return py::make_tuple(code, consts, depth);
}

std::string ols_tree_hook(const std::vector<std::string>& factors, const std::vector<double>& coefs) {
  std::vector<Flat> fs;
  for (const std::string& s : factors) fs.push_back(rrh::parse_prefix(s));
  if (coefs.size() != fs.size() + 1) throw std::invalid_argument("need len(factors)+1 coefficients");
  Flat out;
  rrh::ols_tree(fs, coefs.data(), out);
  return rrh::to_prefix(out);
}

std::vector<int> rand_hook(unsigned seed, int count) {
  rrh::CRand r;
  r.seed(seed);
  std::vector<int> out;
  for (int i = 0; i < count; ++i) out.push_back(r.next());
  return out;
}

// A resident dataset plus a prepared candidate batch: the unit the benchmark
// and the serving loop drive.  prepare() runs the host half of
// tune_constants (expand, simplify, factor split; rils_rols_cpp.cpp:448-475)
// and compiles the factors; run() is one rr_fitness_batch call.
// precision names of the batched runners: "fp64" (RR_FP64), "fp32"
// (RR_FP32_REFINED, uncertified candidates left flagged), "fp32c"
// (RR_FP32_CHECKED, every result final)
int batch_precision(const std::string& p) {
  if (p == "fp64") return RR_FP64;
  if (p == "fp32") return RR_FP32_REFINED;
  if (p == "fp32c") return RR_FP32_CHECKED;
  throw std::invalid_argument("precision must be 'fp64', 'fp32' or 'fp32c'");
}

class BatchRunner {
 public:
  BatchRunner(py::array_t<double, py::array::c_style | py::array::forcecast> X_fm,
              py::array_t<double, py::array::c_style | py::array::forcecast> y, int device) {
    py::buffer_info bx = X_fm.request(), by = y.request();
    if (bx.ndim != 2 || by.ndim != 1 || bx.shape[1] != by.shape[0]) throw std::invalid_argument("X_fm must be (d, n), y (n,)");
    n_ = by.shape[0];
    d_ = (int)bx.shape[0];
    check(rr_create(device, &ctx_), "rr_create");
    check(rr_dataset_upload(ctx_, (const double*)bx.ptr, (const double*)by.ptr, n_, d_, &h_), "rr_dataset_upload");
  }
  ~BatchRunner() { if (ctx_) rr_destroy(ctx_); }

  // OLS mode: returns the number of factors per candidate
  std::vector<int> prepare_ols(const std::vector<std::string>& trees) {
    reset();
    mode_ = RR_MODE_OLS;
    rrh::Pool pool;
    std::vector<Flat> fs;
    std::vector<int> nf;
    for (const std::string& t : trees) {
      rrh::ols_factors(rrh::parse_prefix(t), fs, pool);
      for (const Flat& f : fs) { rrh::compile(f, code_, consts_); prog_.push_back((int32_t)code_.size()); }
      cand_.push_back((int32_t)prog_.size() - 1);
      nf.push_back((int)fs.size());
    }
    return nf;
  }
  // OLS mode from ready factor lists (no rewrites)
  void prepare_factors(const std::vector<std::vector<std::string>>& cands) {
    reset();
    mode_ = RR_MODE_OLS;
    for (const auto& c : cands) {
      for (const std::string& f : c) { rrh::compile(rrh::parse_prefix(f), code_, consts_); prog_.push_back((int32_t)code_.size()); }
      cand_.push_back((int32_t)prog_.size() - 1);
    }
  }
  void prepare_raw(const std::vector<std::string>& trees) {
    reset();
    mode_ = RR_MODE_RAW;
    for (const std::string& t : trees) {
      rrh::compile(rrh::parse_prefix(t), code_, consts_);
      prog_.push_back((int32_t)code_.size());
      cand_.push_back((int32_t)prog_.size() - 1);
    }
  }
  py::tuple run(const std::string& precision) {
    const int pr = batch_precision(precision);
    const int nc = (int)cand_.size() - 1;
    const int64_t ncoef = (int64_t)cand_.back() + nc;
    py::array_t<double> coefs(std::max<int64_t>(ncoef, 1));
    py::array_t<double> fit({(py::ssize_t)std::max(nc, 1), (py::ssize_t)5});
    std::vector<rr_fit> out(std::max(nc, 1));
    rr_batch b{nc, cand_.data(), prog_.data(), code_.data(), consts_.data(), (int32_t)consts_.size()};
    double* cp = (double*)coefs.request().ptr;
    {
      py::gil_scoped_release nogil;
      check(rr_fitness_batch(ctx_, h_, &b, mode_, pr, cp, out.data()), "rr_fitness_batch");
    }
    double* fp = (double*)fit.request().ptr;
    for (int c = 0; c < nc; ++c) {
      fp[5 * c] = out[c].e0; fp[5 * c + 1] = out[c].e1; fp[5 * c + 2] = out[c].err;
      fp[5 * c + 3] = out[c].rank; fp[5 * c + 4] = out[c].flags;
    }
    return py::make_tuple(coefs, fit);
  }
  // k calls of the prepared batch (plain FP64 form), pipelined: call i + 1 is
  // submitted (checked, planned, launched) before call i is waited for, so
  // the host work of one call overlaps the kernels of the previous one.
  // Returns the last call's (coefs, fit), each call's latency from its submit
  // to the end of its wait (s), and the wall time of all k calls (s).
  py::tuple run_pipelined(int k) {
    if (k < 1) throw std::invalid_argument("k must be at least 1");
    const int nc = (int)cand_.size() - 1;
    const int64_t ncoef = (int64_t)cand_.back() + nc;
    std::vector<double> cbuf[2];
    std::vector<rr_fit> fbuf[2];
    for (int i = 0; i < 2; ++i) {
      cbuf[i].resize((size_t)std::max<int64_t>(ncoef, 1));
      fbuf[i].resize((size_t)std::max(nc, 1));
    }
    rr_batch b{nc, cand_.data(), prog_.data(), code_.data(), consts_.data(), (int32_t)consts_.size()};
    std::vector<double> lat((size_t)k);
    std::vector<std::chrono::steady_clock::time_point> ts((size_t)k);
    double total = 0;
    {
      py::gil_scoped_release nogil;
      const auto t0 = std::chrono::steady_clock::now();
      auto since = [](std::chrono::steady_clock::time_point a) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
      };
      for (int i = 0; i < k; ++i) {
        ts[(size_t)i] = std::chrono::steady_clock::now();
        check(rr_fitness_submit(ctx_, h_, &b, mode_, cbuf[i % 2].data(), fbuf[i % 2].data()), "rr_fitness_submit");
        if (i > 0) {
          check(rr_fitness_wait(ctx_), "rr_fitness_wait");
          lat[(size_t)i - 1] = since(ts[(size_t)i - 1]);
        }
      }
      check(rr_fitness_wait(ctx_), "rr_fitness_wait");
      lat[(size_t)k - 1] = since(ts[(size_t)k - 1]);
      total = since(t0);
    }
    const int last = (k - 1) % 2;
    py::array_t<double> coefs(std::max<int64_t>(ncoef, 1));
    py::array_t<double> fit({(py::ssize_t)std::max(nc, 1), (py::ssize_t)5});
    std::copy(cbuf[last].begin(), cbuf[last].end(), (double*)coefs.request().ptr);
    double* fp = (double*)fit.request().ptr;
    for (int c = 0; c < nc; ++c) {
      const rr_fit& f = fbuf[last][(size_t)c];
      fp[5 * c] = f.e0; fp[5 * c + 1] = f.e1; fp[5 * c + 2] = f.err; fp[5 * c + 3] = f.rank; fp[5 * c + 4] = f.flags;
    }
    return py::make_tuple(coefs, fit, lat, total);
  }
  // Sub-batch of already prepared candidates (same constant pool): used to
  // re-score in FP64 the candidates the FP32 form flags.
  void select(const std::vector<int64_t>& idx) {
    scand_.assign(1, 0);
    sprog_.assign(1, 0);
    scode_.clear();
    for (int64_t c : idx) {
      if (c < 0 || c + 1 >= (int64_t)cand_.size()) throw std::out_of_range("candidate index");
      for (int32_t p = cand_[c]; p < cand_[c + 1]; ++p) {
        scode_.insert(scode_.end(), code_.begin() + prog_[p], code_.begin() + prog_[p + 1]);
        sprog_.push_back((int32_t)scode_.size());
      }
      scand_.push_back((int32_t)sprog_.size() - 1);
    }
  }
  py::tuple run_selected(const std::string& precision) {
    const int pr = batch_precision(precision);
    const int nc = (int)scand_.size() - 1;
    const int64_t ncoef = (int64_t)scand_.back() + nc;
    py::array_t<double> coefs(std::max<int64_t>(ncoef, 1));
    py::array_t<double> fit({(py::ssize_t)std::max(nc, 1), (py::ssize_t)5});
    std::vector<rr_fit> out(std::max(nc, 1));
    rr_batch b{nc, scand_.data(), sprog_.data(), scode_.data(), consts_.data(), (int32_t)consts_.size()};
    double* cp = (double*)coefs.request().ptr;
    {
      py::gil_scoped_release nogil;
      check(rr_fitness_batch(ctx_, h_, &b, mode_, pr, cp, out.data()), "rr_fitness_batch");
    }
    double* fp = (double*)fit.request().ptr;
    for (int c = 0; c < nc; ++c) {
      fp[5 * c] = out[c].e0; fp[5 * c + 1] = out[c].e1; fp[5 * c + 2] = out[c].err;
      fp[5 * c + 3] = out[c].rank; fp[5 * c + 4] = out[c].flags;
    }
    return py::make_tuple(coefs, fit);
  }
  double last_kernel_ms() { double ms = 0; rr_last_kernel_ms(ctx_, &ms); return ms; }
  // per-phase kernel time of the last run (option "phase_timing" must be 1)
  py::dict last_phase_ms() {
    static const char* names[RR_PHASES] = {"prepass", "eval", "qr", "score", "fp32", "host_check", "host_plan",
                                           "host_tail"};
    py::dict d;
    for (int p = 0; p < RR_PHASES; ++p) {
      double ms = 0;
      check(rr_last_phase_ms(ctx_, p, &ms), "rr_last_phase_ms");
      d[names[p]] = ms;
    }
    return d;
  }
  // the autotuner on the prepared batch; returns the best kernel time and the chosen options
  py::dict autotune(const std::string& precision, int reps) {
    const int pr = batch_precision(precision);
    rr_batch b{(int32_t)cand_.size() - 1, cand_.data(), prog_.data(), code_.data(), consts_.data(), (int32_t)consts_.size()};
    double ms = 0;
    {
      py::gil_scoped_release nogil;
      check(rr_autotune(ctx_, h_, &b, mode_, pr, reps, &ms), "rr_autotune");
    }
    py::dict d;
    d["best_ms"] = ms;
    for (const char* k : {"fp64_eval_rows", "fp64_eval_form", "fp32_rows", "fp32_chunk_rows", "fp32_streams"}) d[k] = get_option(k);
    return d;
  }
  void set_option(const std::string& k, int64_t v) { check(rr_set_option(ctx_, k.c_str(), v), "rr_set_option"); }
  int64_t get_option(const std::string& k) { int64_t v = 0; check(rr_get_option(ctx_, k.c_str(), &v), "rr_get_option"); return v; }
  std::vector<int> cand_start() const { return std::vector<int>(cand_.begin(), cand_.end()); }
  int64_t code_words() const { return (int64_t)code_.size(); }

 private:
  static void check(int rc, const char* what) { if (rc) throw std::runtime_error(std::string(what) + ": " + rr_last_error()); }
  void reset() { cand_.assign(1, 0); prog_.assign(1, 0); code_.clear(); consts_.clear(); }
  rr_ctx* ctx_ = nullptr;
  int32_t h_ = -1;
  int64_t n_ = 0;
  int d_ = 0;
  int mode_ = RR_MODE_OLS;
  std::vector<int32_t> cand_{0}, prog_{0}, scand_{0}, sprog_{0};
  std::vector<uint32_t> code_, scode_;
  std::vector<double> consts_;
};

}  // namespace

PYBIND11_MODULE(rils_rols_cpp, m) {
  m.doc() = "RILS-ROLS with the candidate-fitness hot path on MI355X (HIP)";
  py::class_<RilsRols>(m, "rils_rols")
      .def(py::init<bool, int, int, double, int, double, bool, int>())
      .def("fit",
           [](RilsRols& self, py::array_t<double, py::array::c_style | py::array::forcecast> X,
              py::array_t<double, py::array::c_style | py::array::forcecast> y, int data_cnt, int feat_cnt) {
             py::buffer_info bx = X.request(), by = y.request();
             const double* xp = (const double*)bx.ptr;
             const double* yp = (const double*)by.ptr;
             py::gil_scoped_release nogil;
             self.fit(xp, bx.size, yp, by.size, data_cnt, feat_cnt);
           })
      .def("local_search",
           [](RilsRols& self, py::array_t<double, py::array::c_style | py::array::forcecast> X,
              py::array_t<double, py::array::c_style | py::array::forcecast> y, int data_cnt, int feat_cnt,
              const std::string& prefix) {
             py::buffer_info bx = X.request(), by = y.request();
             const double* xp = (const double*)bx.ptr;
             const double* yp = (const double*)by.ptr;
             py::gil_scoped_release nogil;
             self.local_search_from(xp, bx.size, yp, by.size, data_cnt, feat_cnt, prefix);
           },
           "One local search (rils_rols_cpp.cpp:601-643) from the tree `prefix` (exchange format); the result "
           "becomes the model")
      .def("predict",
           [](RilsRols& self, py::array_t<double, py::array::c_style | py::array::forcecast> X, int data_cnt,
              int feat_cnt) {
             py::buffer_info bx = X.request();
             std::vector<double> out;
             {
               py::gil_scoped_release nogil;
               out = self.predict((const double*)bx.ptr, bx.size, data_cnt, feat_cnt);
             }
             py::array_t<double> res(data_cnt);
             std::copy(out.begin(), out.end(), (double*)res.request().ptr);
             return res;
           })
      .def("get_model_string", &RilsRols::get_model_string)
      .def("get_best_time", &RilsRols::get_best_time)
      .def("get_fit_calls", &RilsRols::get_fit_calls)
      .def("get_total_time", &RilsRols::get_total_time)
      // beyond the reference
      .def("get_moves", [](const RilsRols& self) { return moves_list(self.moves()); })
      .def("get_sweep_ms", &RilsRols::sweep_ms)
      .def("get_stats", [](const RilsRols& self) { return stats_dict(self.stats()); })
      .def("get_model_prefix", &RilsRols::final_prefix)
      .def("set_precision",
           [](RilsRols& self, const std::string& p) {
             if (p == "fp64") self.set_precision(RR_FP64);
             else if (p == "fp32") self.set_precision(RR_FP32_REFINED);
             else throw std::invalid_argument("precision must be 'fp64' or 'fp32'");
           })
      .def("set_device", &RilsRols::set_device)
      .def("set_slab_budget_mb", &RilsRols::set_slab_budget_mb)
      .def("set_context_option", &RilsRols::set_context_option, py::arg("name"), py::arg("value"))
      .def("set_autotune", &RilsRols::set_autotune)
      .def("set_host_threads", &RilsRols::set_host_threads,
           "host threads for the per-candidate rewrites of a sweep (0: min(hardware, 8))");

  py::class_<BatchRunner>(m, "BatchRunner")
      .def(py::init<py::array_t<double, py::array::c_style | py::array::forcecast>,
                    py::array_t<double, py::array::c_style | py::array::forcecast>, int>(),
           py::arg("X_fm"), py::arg("y"), py::arg("device") = 0)
      .def("prepare_ols", &BatchRunner::prepare_ols)
      .def("prepare_factors", &BatchRunner::prepare_factors)
      .def("prepare_raw", &BatchRunner::prepare_raw)
      .def("run", &BatchRunner::run, py::arg("precision") = "fp64")
      .def("run_pipelined", &BatchRunner::run_pipelined, py::arg("k"))
      .def("select", &BatchRunner::select)
      .def("run_selected", &BatchRunner::run_selected, py::arg("precision") = "fp64")
      .def("last_kernel_ms", &BatchRunner::last_kernel_ms)
      .def("last_phase_ms", &BatchRunner::last_phase_ms)
      .def("set_option", &BatchRunner::set_option)
      .def("autotune", &BatchRunner::autotune, py::arg("precision"), py::arg("reps") = 3)
      .def("get_option", &BatchRunner::get_option)
      .def("cand_start", &BatchRunner::cand_start)
      .def("code_words", &BatchRunner::code_words);

  // test hooks (host engine only; no GPU needed)
  py::class_<rrh::Server>(m, "Server",
                           "Request loop: many resident RILS-ROLS runs on one GPU, their fitness work merged into "
                           "shared launches as it arrives, runs admitted by HBM footprint")
      .def(py::init<int, int64_t, int, int, int64_t, int>(), py::arg("device") = 0, py::arg("hbm_budget_mb") = 0,
           py::arg("max_active") = 64, py::arg("host_threads_per_run") = 1, py::arg("slab_budget_mb") = 0,
           py::arg("gather_us") = 0)
      .def(
          "submit",
          [](rrh::Server& self, py::array_t<double, py::array::c_style | py::array::forcecast> X,
             py::array_t<double, py::array::c_style | py::array::forcecast> y, bool classification, int max_fit_calls,
             int max_seconds, double complexity_penalty, int max_complexity, double sample_size, int random_state,
             const std::string& precision) {
            py::buffer_info bx = X.request(), by = y.request();
            if (bx.ndim != 2) throw std::invalid_argument("X must be a 2-d array");
            rrh::RunSpec r;
            r.n = (int)bx.shape[0];
            r.d = (int)bx.shape[1];
            r.X.assign((const double*)bx.ptr, (const double*)bx.ptr + bx.size);
            r.y.assign((const double*)by.ptr, (const double*)by.ptr + by.size);
            r.classification = classification;
            r.max_fit_calls = max_fit_calls;
            r.max_seconds = max_seconds;
            r.complexity_penalty = complexity_penalty;
            r.max_complexity = max_complexity;
            r.sample_size = sample_size;
            r.random_state = random_state;
            if (precision == "fp64") r.precision = RR_FP64;
            else if (precision == "fp32") r.precision = RR_FP32_REFINED;
            else throw std::invalid_argument("precision must be 'fp64' or 'fp32'");
            return self.submit(std::move(r));
          },
          py::arg("X"), py::arg("y"), py::arg("classification") = false, py::arg("max_fit_calls") = 100000,
          py::arg("max_seconds") = 100, py::arg("complexity_penalty") = 0.001, py::arg("max_complexity") = 50,
          py::arg("sample_size") = 1.0, py::arg("random_state") = 0, py::arg("precision") = "fp64")
      .def("done", &rrh::Server::done)
      .def("wait",
           [](rrh::Server& self, int id) {
             rrh::RunResult r;
             {
               py::gil_scoped_release nogil;
               r = self.wait(id);
             }
             py::dict d;
             d["id"] = r.id; d["ok"] = r.ok; d["error"] = r.error; d["model"] = r.model; d["prefix"] = r.prefix;
             d["fit_calls"] = r.fit_calls; d["moves"] = moves_list(r.moves); d["sweep_ms"] = r.sweep_ms;
             d["stats"] = stats_dict(r.stats); d["submitted_s"] = r.submitted_s; d["admitted_s"] = r.admitted_s;
             d["finished_s"] = r.finished_s; d["footprint"] = r.footprint;
             return d;
           })
      .def("resident_bytes", &rrh::Server::resident_bytes)
      .def("stats", [](rrh::Server& self) {
        const rrh::Dispatcher::Counters c = self.dispatcher_counters();
        py::dict d;
        d["launches"] = c.launches; d["requests"] = c.requests; d["candidates"] = c.candidates;
        d["max_requests_per_launch"] = c.max_requests_per_launch; d["kernel_ms"] = c.kernel_ms; d["busy_s"] = c.busy_s;
        d["gather_s"] = c.gather_s;
        return d;
      })
      .def_static("footprint", &rrh::Server::footprint);
  m.def("_merge_batches", &merge_hook);
  m.def("_all_candidates", &candidates_hook, py::arg("prefix"), py::arg("nvars"), py::arg("classification") = false,
        py::arg("local_search") = true);
  m.def("_rewrite", &rewrite_hook);
  m.def("_relevant_features", &features_hook);
  m.def("_ols_factors", &factors_hook);
  m.def("_compile", &compile_hook);
  m.def("_ols_tree", &ols_tree_hook);
  m.def("_crand", &rand_hook);
}
