// extern "C" boundary of libflygpu (declared in include/flygpu.h).  Every entry point catches C++
// exceptions and returns a status code; the message is kept on the context.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>

#include "perturb.hpp"
#include "plain.hpp"

namespace flyg {
  void list_positions(flyg_list_s& nl, double* d_out);
  long list_count(flyg_list_s& nl, double rc, std::vector<int>& counts);
  void list_export(flyg_list_s& nl, double rc, int64_t* offsets, int32_t* j, int32_t* shift, double* dr);
}  // namespace flyg

using namespace flyg;

void flyg_ctx_s::set_device() const { HIPCHK(hipSetDevice(device)); }

void flyg_ctx_s::release() {
  if (--refs > 0) return;
  (void)hipSetDevice(device);
  (void)hipStreamSynchronize(stream);
  if (hflags) (void)hipHostFree(hflags);
  if (hscal) (void)hipHostFree(hscal);
  (void)hipStreamDestroy(stream);
  delete this;   // device buffers are freed by their destructors
}

unsigned flyg_ctx_s::take_flags() {
  HIPCHK(hipMemcpyAsync(hflags, dflags.get(), sizeof(unsigned), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  unsigned f = hflags[0];
  if (f) HIPCHK(hipMemsetAsync(dflags.get(), 0, sizeof(unsigned), stream));
  return f;
}

namespace {
  thread_local std::string g_err;

  template <typename F>
  int guard(flyg_ctx_s* ctx, F&& f) {
    try {
      if (ctx) ctx->set_device();
      f();
      return FLYG_OK;
    } catch (Error const& e) {
      (ctx ? ctx->err : g_err) = e.what();
      return e.status;
    } catch (std::exception const& e) {
      (ctx ? ctx->err : g_err) = e.what();
      return FLYG_INVALID_ARG;
    }
  }

  // Copies a host or device vector of `count` T into a device buffer (or returns the device pointer).
  template <typename T>
  T const* in_dev(flyg_ctx_s& ctx, T const* p, std::size_t count, int where, DBuf<T>& stage) {
    if (where == FLYG_DEVICE) return p;
    require(where == FLYG_HOST, "`where` must be FLYG_HOST or FLYG_DEVICE");
    stage.reserve(count);
    HIPCHK(hipMemcpyAsync(stage.get(), p, count * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
    return stage.get();
  }

  void check_errors(flyg_ctx_s& ctx) {
    unsigned fl = ctx.take_flags();
    if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
    if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
    if (fl) throw Error(FLYG_DEVICE_ERROR, "device error flag set");
  }
}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------------------

int flyg_ctx_create(int device, flyg_ctx* out) {
  return guard(nullptr, [&] {
    require(out != nullptr, "null output handle");
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "no such HIP device");
    HIPCHK(hipSetDevice(device));
    auto c = std::make_unique<flyg_ctx_s>();
    c->device = device;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    noalloc_scratch_init(device);   // the allocation-free gradient's scratch address on this device
    c->dflags.reserve(16);
    HIPCHK(hipMemset(c->dflags.get(), 0, 16 * sizeof(unsigned)));
    c->dscal.reserve(64);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hflags), 16 * sizeof(unsigned), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->hscal), 64 * sizeof(double), hipHostMallocDefault));
    *out = c.release();
  });
}

int flyg_ctx_destroy(flyg_ctx ctx) {
  if (ctx) ctx->release();
  return FLYG_OK;
}

const char* flyg_ctx_last_error(flyg_ctx ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int flyg_ctx_synchronize(flyg_ctx ctx) {
  return guard(ctx, [&] { check_errors(*ctx); });
}

void* flyg_ctx_stream(flyg_ctx ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int flyg_ctx_device_info(flyg_ctx ctx, char* name, size_t len, int* cus) {
  return guard(ctx, [&] {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    if (name && len) {
      std::strncpy(name, prop.gcnArchName, len - 1);
      name[len - 1] = 0;
    }
    if (cus) *cus = prop.multiProcessorCount;
  });
}

// ---------------------------------------------------------------------------------------------------
// Potential
// ---------------------------------------------------------------------------------------------------

int flyg_eam_load(flyg_ctx ctx, const char* path, flyg_eam* out) {
  return guard(ctx, [&] {
    require(ctx && path && out, "null argument");
    auto e = std::make_unique<flyg_eam_s>();
    e->ctx = ctx;
    e->data = read_setfl(path);
    eam_setup_tables(*e);
    ctx->retain();
    *out = e.release();
  });
}

int flyg_eam_load_text(flyg_ctx ctx, const char* text, size_t len, flyg_eam* out) {
  return guard(ctx, [&] {
    require(ctx && text && out, "null argument");
    auto e = std::make_unique<flyg_eam_s>();
    e->ctx = ctx;
    std::istringstream in(std::string(text, len));
    e->data = read_setfl_stream(in);
    eam_setup_tables(*e);
    ctx->retain();
    *out = e.release();
  });
}

int flyg_eam_destroy(flyg_eam eam) {
  if (!eam) return FLYG_OK;
  flyg_ctx_s* ctx = eam->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete eam;
  ctx->release();
  return FLYG_OK;
}

int flyg_eam_info(flyg_eam eam, int* ntypes, double* r_cut, double* masses, char* names) {
  return guard(eam->ctx, [&] {
    if (ntypes) *ntypes = eam->data.ntypes;
    if (r_cut) *r_cut = eam->data.rcut;
    for (int i = 0; i < eam->data.ntypes; ++i) {
      if (masses) masses[i] = eam->data.mass[i];
      if (names) {
        std::memset(names + 16 * i, 0, 16);
        std::strncpy(names + 16 * i, eam->data.names[i].c_str(), 15);
      }
    }
  });
}

int flyg_eam_set_precision(flyg_eam eam, int precision) {
  return guard(eam->ctx, [&] {
    require(precision == FLYG_FP64 || precision == FLYG_MIXED, "precision must be FLYG_FP64 or FLYG_MIXED");
    eam->precision = precision;
  });
}

int flyg_eam_set_lanes(flyg_eam eam, int lanes) {
  return guard(eam->ctx, [&] {
    require(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16, "lanes must be 0, 4, 8 or 16");
    eam->lanes = lanes;
  });
}

int flyg_eam_set_timing(flyg_eam eam, int on) {
  return guard(eam->ctx, [&] {
    if (on) eam->events_used = 0;   // a new record; turning timing off keeps the last one readable
    eam->timing = on != 0;
  });
}

int flyg_eam_timing(flyg_eam eam, double* ms_total, int64_t* calls) {
  return guard(eam->ctx, [&] {
    require(ms_total && calls, "null argument");
    HIPCHK(hipStreamSynchronize(eam->ctx->stream));
    double t = 0;
    for (std::size_t k = 0; k < eam->events_used; ++k) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, eam->events[k].first, eam->events[k].second));
      t += ms;
    }
    *ms_total = t;
    *calls = static_cast<int64_t>(eam->events_used);
  });
}

int flyg_eam_table_eval(flyg_eam eam, int kind, int a, int b, const double* x, int64_t m, int deriv, double* out) {
  return guard(eam->ctx, [&] {
    require(kind >= 0 && kind <= 2, "kind must be 0 (F), 1 (phi) or 2 (V)");
    require(m >= 0, "negative count");
    if (m == 0) return;
    eam_table_eval(*eam, kind, a, b, x, m, deriv, out);
  });
}

// ---------------------------------------------------------------------------------------------------
// Neighbour list
// ---------------------------------------------------------------------------------------------------

int flyg_list_create(flyg_ctx ctx, const double basis[9], const int periodic[3], double r_cut, flyg_list* out) {
  return guard(ctx, [&] {
    require(ctx && basis && periodic && out, "null argument");
    *out = make_list(ctx, basis, periodic, r_cut).release();
    ctx->retain();
  });
}

int flyg_list_destroy(flyg_list nl) {
  if (!nl) return FLYG_OK;
  flyg_ctx_s* ctx = nl->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete nl;
  ctx->release();
  return FLYG_OK;
}

int flyg_list_rebuild(flyg_list nl, const double* pos, int64_t n, int where) {
  return guard(nl->ctx, [&] {
    require(n > 0, "List::rebuild with no atoms");
    double const* d = in_dev(*nl->ctx, pos, 3 * static_cast<std::size_t>(n), where, nl->stage);
    list_rebuild(*nl, d, n);
  });
}

int flyg_list_update(flyg_list nl, const double* delta, int64_t n, int where) {
  return guard(nl->ctx, [&] {
    require(n == nl->n, "Input to update has the wrong number of atoms");
    double const* d = in_dev(*nl->ctx, delta, 3 * static_cast<std::size_t>(n), where, nl->stage);
    list_update(*nl, d, n);
    if (where == FLYG_HOST) HIPCHK(hipStreamSynchronize(nl->ctx->stream));
  });
}

int flyg_list_size(flyg_list nl, int64_t* n) {
  return guard(nl->ctx, [&] { *n = nl->n; });
}

int flyg_list_num_ghosts(flyg_list nl, int64_t* g) {
  return guard(nl->ctx, [&] { *g = nl->nghost; });
}

int flyg_list_count(flyg_list nl, double r_cut, int64_t* n_pairs) {
  return guard(nl->ctx, [&] {
    std::vector<int> counts;
    *n_pairs = list_count(*nl, r_cut, counts);
  });
}

int flyg_list_export(flyg_list nl, double r_cut, int64_t* offsets, int32_t* j, int32_t* shift, double* dr) {
  return guard(nl->ctx, [&] { list_export(*nl, r_cut, offsets, j, shift, dr); });
}

int flyg_list_positions(flyg_list nl, double* pos, int where) {
  return guard(nl->ctx, [&] {
    require(nl->built, "List used before rebuild");
    if (where == FLYG_DEVICE) {
      list_positions(*nl, pos);
      return;
    }
    nl->stage.reserve(3 * nl->n);
    list_positions(*nl, nl->stage.get());
    HIPCHK(hipMemcpyAsync(pos, nl->stage.get(), 3 * nl->n * sizeof(double), hipMemcpyDeviceToHost, nl->ctx->stream));
    HIPCHK(hipStreamSynchronize(nl->ctx->stream));
  });
}

// ---------------------------------------------------------------------------------------------------
// EAM evaluation
// ---------------------------------------------------------------------------------------------------

int flyg_eam_energy(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n, int where,
                    double* energy) {
  return guard(eam->ctx, [&] {
    require(nl->ctx == eam->ctx, "list and potential belong to different contexts");
    require(n == nl->n, "EAM energy size mismatch");
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    *energy = eam_energy(*eam, *nl, t, f);
  });
}

int flyg_eam_stage_atoms(flyg_eam eam, const uint32_t* types, const uint8_t* frozen, int64_t n, int where) {
  return guard(eam->ctx, [&] {
    require(n > 0, "no atoms");
    require(types && frozen, "null argument");
    if (where == FLYG_DEVICE) {
      stage_atoms_device(*eam, types, frozen, n);
    } else {
      require(where == FLYG_HOST, "`where` must be FLYG_HOST or FLYG_DEVICE");
      uint32_t const* t;
      uint8_t const* f;
      stage_atoms(*eam, types, frozen, n, FLYG_HOST, &t, &f);
    }
  });
}

int flyg_eam_gradient(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n, int where,
                      double* grad) {
  return guard(eam->ctx, [&] {
    require(nl->ctx == eam->ctx, "list and potential belong to different contexts");
    require(n == nl->n, "EAM gradient size mismatch");
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    if (where == FLYG_DEVICE) {
      eam_gradient(*eam, *nl, t, f, grad, eam->precision);   // asynchronous; errors at the next sync
      return;
    }
    eam->vec_out.reserve(3 * n);
    eam_gradient(*eam, *nl, t, f, eam->vec_out.get(), eam->precision);
    HIPCHK(hipMemcpyAsync(grad, eam->vec_out.get(), 3 * n * sizeof(double), hipMemcpyDeviceToHost, eam->ctx->stream));
    check_errors(*eam->ctx);
  });
}

int flyg_eam_gradient_noalloc(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n,
                              double* grad) {
  return guard(eam->ctx, [&] {
    require(nl->ctx == eam->ctx, "list and potential belong to different contexts");
    require(n == nl->n, "EAM gradient size mismatch");
    require(types && frozen && grad, "null argument");
    eam_gradient_noalloc(*eam, *nl, types, frozen, grad, eam->precision);   // asynchronous; errors at the next sync
  });
}

int flyg_eam_hessian_eigenvalues(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n,
                                 int where, double* eigenvalues, int64_t* n_dof) {
  return guard(eam->ctx, [&] {
    require(nl->ctx == eam->ctx, "list and potential belong to different contexts");
    require(n == nl->n, "EAM Hessian size mismatch");
    require(types && frozen && eigenvalues && n_dof, "null argument");
    std::vector<uint32_t> ht(types, types + (where == FLYG_HOST ? n : 0));
    std::vector<uint8_t> hf(frozen, frozen + (where == FLYG_HOST ? n : 0));
    if (where == FLYG_DEVICE) {
      ht.resize(n);
      hf.resize(n);
      HIPCHK(hipMemcpyAsync(ht.data(), types, n * sizeof(uint32_t), hipMemcpyDeviceToHost, eam->ctx->stream));
      HIPCHK(hipMemcpyAsync(hf.data(), frozen, n * sizeof(uint8_t), hipMemcpyDeviceToHost, eam->ctx->stream));
      HIPCHK(hipStreamSynchronize(eam->ctx->stream));
    }
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    std::vector<double> w = hessian_eigenvalues(*eam, *nl, t, f, hf, ht, nullptr);
    std::copy(w.begin(), w.end(), eigenvalues);
    *n_dof = static_cast<int64_t>(w.size());
  });
}

int flyg_log_prod_eigen(const double* eigenvalues, int64_t n, double zero_tol, double* log_prod, int64_t* n_zero,
                        int64_t* n_negative) {
  return guard(nullptr, [&] {
    require(eigenvalues && log_prod && n_zero && n_negative && n >= 0, "null argument");
    double lp = 0;
    int64_t z = 0, neg = 0;
    for (int64_t i = 0; i < n; ++i) {
      double e = eigenvalues[i];
      if (e < -zero_tol) {
        ++neg;
      } else if (e > zero_tol) {
        lp += std::log(e);
      } else {
        ++z;
      }
    }
    *log_prod = lp;
    *n_zero = z;
    *n_negative = neg;
  });
}

int flyg_prefactor(double log_prod_min, double log_prod_sp, double* prefactor) {
  return guard(nullptr, [&] {
    require(prefactor != nullptr, "null argument");
    *prefactor = std::sqrt(std::exp(log_prod_min - log_prod_sp) / (2 * M_PI * 1.6605390666050e-27));
  });
}

int flyg_eam_hvp(flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen, int64_t n, const double* v,
                 double* hv, int where) {
  return guard(eam->ctx, [&] {
    require(nl->ctx == eam->ctx, "list and potential belong to different contexts");
    require(n == nl->n, "EAM hvp size mismatch");
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    if (where == FLYG_DEVICE) {
      eam_hvp(*eam, *nl, t, f, v, hv);
      return;
    }
    double const* dv = in_dev(*eam->ctx, v, 3 * n, where, eam->vec_in);
    eam->vec_out.reserve(3 * n);
    eam_hvp(*eam, *nl, t, f, dv, eam->vec_out.get());
    HIPCHK(hipMemcpyAsync(hv, eam->vec_out.get(), 3 * n * sizeof(double), hipMemcpyDeviceToHost, eam->ctx->stream));
    check_errors(*eam->ctx);
  });
}

// ---------------------------------------------------------------------------------------------------
// StepLBFGS
// ---------------------------------------------------------------------------------------------------

int flyg_steplbfgs_create(flyg_ctx ctx, int m, flyg_steplbfgs* out) {
  return guard(ctx, [&] {
    auto s = std::make_unique<flyg_steplbfgs_s>();
    s->ctx = ctx;
    s->core = std::make_unique<DevLBFGS>(ctx, m);
    ctx->retain();
    *out = s.release();
  });
}

int flyg_steplbfgs_destroy(flyg_steplbfgs s) {
  if (!s) return FLYG_OK;
  flyg_ctx_s* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete s;
  ctx->release();
  return FLYG_OK;
}

int flyg_steplbfgs_clear(flyg_steplbfgs s) {
  return guard(s->ctx, [&] { s->core->clear(); });
}

int flyg_steplbfgs_newton_step(flyg_steplbfgs s, const double* x, const double* g, int64_t len, double* q, int where) {
  return guard(s->ctx, [&] {
    require(len > 0, "empty newton_step");
    flyg_ctx_s& ctx = *s->ctx;
    double const* dx = x;
    double const* dg = g;
    if (where == FLYG_HOST) {
      HIPCHK(hipMemcpyAsync(s->x.p(len), x, len * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      HIPCHK(hipMemcpyAsync(s->g.p(len), g, len * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      dx = s->x.get();
      dg = s->g.get();
    }
    double* r = s->core->newton_step(dx, dg, len);
    HIPCHK(hipMemcpyAsync(q, r, len * sizeof(double), where == FLYG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                          ctx.stream));
    if (where == FLYG_HOST) HIPCHK(hipStreamSynchronize(ctx.stream));
  });
}

// ---------------------------------------------------------------------------------------------------
// LBFGS minimiser
// ---------------------------------------------------------------------------------------------------

void flyg_lbfgs_default_options(flyg_lbfgs_options* o) {
  *o = flyg_lbfgs_options{10, 2000, 1e-6, 1.1, 0.0, 0.5, 0.01, 1.5, 0.5};
}

int flyg_lbfgs_create(flyg_ctx ctx, const flyg_lbfgs_options* opt, const double basis[9], const int periodic[3],
                      flyg_lbfgs* out) {
  return guard(ctx, [&] {
    auto m = std::make_unique<flyg_lbfgs_s>();
    m->ctx = ctx;
    if (opt) {
      m->opt = *opt;
    } else {
      flyg_lbfgs_default_options(&m->opt);
    }
    require(m->opt.n > 0 && m->opt.n <= 32, "L-BFGS history length must be in [1, 32]");
    m->box = make_box(basis, periodic);
    std::memcpy(m->basis, basis, sizeof(m->basis));
    std::memcpy(m->periodic, periodic, sizeof(m->periodic));
    ctx->retain();
    *out = m.release();
  });
}

int flyg_lbfgs_destroy(flyg_lbfgs m) {
  if (!m) return FLYG_OK;
  flyg_ctx_s* ctx = m->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete m;
  ctx->release();
  return FLYG_OK;
}

int flyg_lbfgs_minimise(flyg_lbfgs m, flyg_eam eam, double* pos, double* grad, const uint32_t* types,
                        const uint8_t* frozen, int64_t n, int where, int* failed, int* iterations) {
  return guard(m->ctx, [&] {
    require(eam->ctx == m->ctx, "minimiser and potential belong to different contexts");
    require(n > 0, "no atoms");
    flyg_ctx_s& ctx = *m->ctx;
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    long N = 3 * n;
    double* x = pos;
    double* g = grad;
    if (where == FLYG_HOST) {
      x = m->x.p(N);
      g = m->g.p(N);
      HIPCHK(hipMemcpyAsync(x, pos, N * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    }
    int it = 0;
    bool fail = lbfgs_minimise(*m, *eam, x, g, t, f, n, &it);
    if (where == FLYG_HOST) {
      HIPCHK(hipMemcpyAsync(pos, x, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
      HIPCHK(hipMemcpyAsync(grad, g, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    }
    check_errors(ctx);
    if (failed) *failed = fail ? 1 : 0;
    if (iterations) *iterations = it;
  });
}

int flyg_lbfgs_set_observer(flyg_lbfgs m, flyg_step_observer fn, void* user) {
  return guard(m->ctx, [&] {
    m->obs = fn;
    m->obs_user = user;
  });
}

// ---------------------------------------------------------------------------------------------------
// Dimer
// ---------------------------------------------------------------------------------------------------

void flyg_dimer_default_options(flyg_dimer_options* o) {
  *o = flyg_dimer_options{10, 1000, 5, 1e-5, 1.1, 0.0, 0.5, 0.05, 1.5, 0.5, 3, 1};
}

void flyg_rotor_default_options(flyg_rotor_options* o) { *o = flyg_rotor_options{6, 20, 0.001, M_PI / 360., 1}; }

int flyg_dimer_create(flyg_ctx ctx, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                      const double basis[9], const int periodic[3], flyg_dimer* out) {
  return guard(ctx, [&] {
    auto d = std::make_unique<flyg_dimer_s>();
    d->ctx = ctx;
    if (opt) {
      d->opt = *opt;
    } else {
      flyg_dimer_default_options(&d->opt);
    }
    if (ropt) {
      d->ropt = *ropt;
    } else {
      flyg_rotor_default_options(&d->ropt);
    }
    require(d->opt.n > 0 && d->opt.n <= 32 && d->ropt.n > 0 && d->ropt.n <= 32, "L-BFGS history length must be in [1, 32]");
    require(d->opt.hist_check_freq > 0, "hist_check_freq must be positive");
    (void)make_box(basis, periodic);
    std::memcpy(d->basis, basis, sizeof(d->basis));
    std::memcpy(d->periodic, periodic, sizeof(d->periodic));
    ctx->retain();
    *out = d.release();
  });
}

int flyg_dimer_destroy(flyg_dimer d) {
  if (!d) return FLYG_OK;
  flyg_ctx_s* ctx = d->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete d;
  ctx->release();
  return FLYG_OK;
}

int flyg_rotor_create(flyg_ctx ctx, const flyg_rotor_options* opt, flyg_rotor* out) {
  return guard(ctx, [&] {
    require(out != nullptr, "null output handle");
    auto r = std::make_unique<flyg_rotor_s>();
    r->ctx = ctx;
    r->d.ctx = ctx;
    if (opt) {
      r->d.ropt = *opt;
    } else {
      flyg_rotor_default_options(&r->d.ropt);
    }
    require(r->d.ropt.n > 0 && r->d.ropt.n <= 32, "L-BFGS history length must be in [1, 32]");
    require(r->d.ropt.delta_r > 0, "delta_r must be positive");
    ctx->retain();
    *out = r.release();
  });
}

int flyg_rotor_destroy(flyg_rotor r) {
  if (!r) return FLYG_OK;
  flyg_ctx_s* ctx = r->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete r;
  ctx->release();
  return FLYG_OK;
}

int flyg_rotor_r_cut(flyg_rotor r, flyg_eam eam, double* r_cut) {
  return guard(r->ctx, [&] {
    require(r_cut != nullptr, "null argument");
    *r_cut = eam->tab.rc + r->d.ropt.delta_r * 2;
  });
}

int flyg_rotor_eff_gradient(flyg_rotor r, flyg_eam eam, flyg_list nl, const uint32_t* types, const uint8_t* frozen,
                            int64_t n, double* axis, double* eff, int count_frozen, int where, double* curv) {
  return guard(r->ctx, [&] {
    require(eam->ctx == r->ctx && nl->ctx == r->ctx, "rotor, potential and list belong to different contexts");
    require(n > 0 && axis && eff && curv, "null or empty argument");
    flyg_ctx_s& ctx = *r->ctx;
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    long N = 3 * n;
    double* ax = axis;
    double* ef = eff;
    flyg::DVec hax, hef;
    if (where == FLYG_HOST) {
      ax = hax.p(N);
      ef = hef.p(N);
      HIPCHK(hipMemcpyAsync(ax, axis, N * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
    }
    double c = rotor_eff_gradient(r->d, *nl, *eam, t, f, n, ax, ef, count_frozen);
    if (where == FLYG_HOST) {
      HIPCHK(hipMemcpyAsync(axis, ax, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
      HIPCHK(hipMemcpyAsync(eff, ef, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    }
    check_errors(ctx);
    *curv = c;
  });
}

int flyg_dimer_find_sp(flyg_dimer d, flyg_eam eam, double* pos, double* axis, const uint32_t* types,
                       const uint8_t* frozen, int64_t n, const double* in_min, const double* hist_sp, int n_hist,
                       double theta_tol, int count_frozen, int where, flyg_search_stats* stats) {
  return guard(d->ctx, [&] {
    require(eam->ctx == d->ctx, "dimer and potential belong to different contexts");
    require(n > 0 && n_hist >= 0, "bad sizes");
    require(n_hist == 0 || hist_sp != nullptr, "hist_sp is null");
    flyg_ctx_s& ctx = *d->ctx;
    uint32_t const* t;
    uint8_t const* f;
    stage_atoms(*eam, types, frozen, n, where, &t, &f);
    long N = 3 * n;
    double* x = pos;
    double* ax = axis;
    double const* mn = in_min;
    double const* hs = hist_sp;
    flyg::DVec hx, hax, hmn, hh;
    if (where == FLYG_HOST) {
      x = hx.p(N);
      ax = hax.p(N);
      HIPCHK(hipMemcpyAsync(x, pos, N * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      HIPCHK(hipMemcpyAsync(ax, axis, N * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      double* m = hmn.p(N);
      HIPCHK(hipMemcpyAsync(m, in_min, N * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
      mn = m;
      if (n_hist > 0) {
        double* h = hh.p(N * n_hist);
        HIPCHK(hipMemcpyAsync(h, hist_sp, N * n_hist * sizeof(double), hipMemcpyHostToDevice, ctx.stream));
        hs = h;
      }
    }
    flyg_search_stats st{};
    dimer_find_sp(*d, *eam, x, ax, t, f, n, mn, hs, n_hist, theta_tol, count_frozen, &st);
    if (where == FLYG_HOST) {
      HIPCHK(hipMemcpyAsync(pos, x, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
      HIPCHK(hipMemcpyAsync(axis, ax, N * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    }
    check_errors(ctx);
    if (stats) *stats = st;
  });
}

int flyg_dimer_set_observer(flyg_dimer d, flyg_step_observer fn, void* user) {
  return guard(d->ctx, [&] {
    d->obs = fn;
    d->obs_user = user;
  });
}

// ---------------------------------------------------------------------------------------------------
// Master::perturb
// ---------------------------------------------------------------------------------------------------

int flyg_perturb(const double basis[9], const int periodic[3], const double* pos, const uint8_t* frozen, int64_t n,
                 int64_t centre, double stddev, double r_pert, uint64_t prng[4], double* out_pos, double* out_axis,
                 double* sd) {
  return guard(nullptr, [&] {
    require(basis && periodic && pos && frozen && prng && out_pos && out_axis && n > 0, "null or empty argument");
    BoxDesc box = make_box(basis, periodic);
    Xoshiro256 g{{prng[0], prng[1], prng[2], prng[3]}};
    Kick k;
    perturb(box, pos, n, centre, nullptr, [&](long j) { return frozen[j] != 0; }, stddev, r_pert, g, k);
    std::copy(pos, pos + 3 * n, out_pos);
    std::fill(out_axis, out_axis + 3 * n, 0.0);
    for (std::size_t q = 0; q < k.atoms.size(); ++q)
      for (int c = 0; c < 3; ++c) {
        out_pos[3 * k.atoms[q] + c] = k.pos[3 * q + c];
        out_axis[3 * k.atoms[q] + c] = k.axis[3 * q + c];
      }
    for (int c = 0; c < 4; ++c) prng[c] = g.s[c];
    if (sd) *sd = k.sd;
  });
}

int flyg_theta_tol(int tot, double* theta_tol) {
  return guard(nullptr, [&] {
    require(theta_tol != nullptr, "null argument");
    *theta_tol = ((30 - 2.5) * std::exp(-0.02 * tot) + 2.5) / 360. * 2. * M_PI;
  });
}

}  // extern "C"
