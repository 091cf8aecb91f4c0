// Header-only C++ drop-in for the reference's saddle-search hot-path API (namespaces, class names and
// member signatures of fly::neigh::List, fly::potential::{DataEAM, EAM, Generic},
// fly::minimise::{StepLBFGS, LBFGS}, fly::saddle::{Rotor, Dimer} and the SoA/property/Box types they
// take), implemented over the C-ABI of libflygpu (include/flygpu.h).  No Eigen: the small value types
// (Vec, Mat, Arr) and the SoA container are re-implemented for the subset the hot path uses.
//
// Errors: status codes from the C-ABI are re-thrown as fly::RuntimeError or fly::Spline::OutOfBounds,
// so reference-style callers (Master::find_batch's retry on OutOfBounds, find.cpp:706-723) behave the
// same.  The `threads` arguments are accepted and ignored (the work runs on the GPU).
// Each host thread uses its own device context (the reference keeps per-thread state in Master's
// ThreadData, find.hpp:310-317).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iterator>
#include <memory>
#include <numeric>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <random>
#include <utility>
#include <vector>

#include "../../flygpu.h"

namespace fly {

  // ---- errors (core.hpp:89-171, spline.hpp:46-51) ------------------------------------------------
  struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
  };

  struct Spline {
    struct OutOfBounds : RuntimeError {
      using RuntimeError::RuntimeError;
    };
  };

  namespace gpu {
    // Per-thread device context.
    inline flyg_ctx context() {
      struct Holder {
        flyg_ctx h = nullptr;
        ~Holder() {
          if (h) flyg_ctx_destroy(h);
        }
      };
      thread_local Holder holder;
      if (!holder.h) {
        int dev = 0;
        if (char const* e = std::getenv("FLYG_DEVICE")) dev = std::atoi(e);
        if (flyg_ctx_create(dev, &holder.h) != FLYG_OK) throw RuntimeError(flyg_ctx_last_error(nullptr));
      }
      return holder.h;
    }

    inline void check(int rc, flyg_ctx ctx = nullptr) {
      if (rc == FLYG_OK) return;
      std::string msg = flyg_ctx_last_error(ctx ? ctx : context());
      if (rc == FLYG_SPLINE_OOB) throw Spline::OutOfBounds(msg);
      throw RuntimeError(msg);
    }
  }  // namespace gpu

  // ---- small value types ---------------------------------------------------------------------------
  constexpr int spatial_dims = 3;

  struct Vec {
    std::array<double, 3> v{0, 0, 0};
    Vec() = default;
    Vec(double x, double y, double z) : v{x, y, z} {}
    static Vec Zero() { return {}; }
    static Vec Constant(double c) { return {c, c, c}; }
    double& operator[](int i) { return v[static_cast<std::size_t>(i)]; }
    double operator[](int i) const { return v[static_cast<std::size_t>(i)]; }
    double& operator()(int i) { return v[static_cast<std::size_t>(i)]; }
    double operator()(int i) const { return v[static_cast<std::size_t>(i)]; }
    friend Vec operator+(Vec a, Vec const& b) {
      for (int k = 0; k < 3; ++k) a[k] += b[k];
      return a;
    }
    friend Vec operator-(Vec a, Vec const& b) {
      for (int k = 0; k < 3; ++k) a[k] -= b[k];
      return a;
    }
    friend Vec operator*(double s, Vec a) {
      for (int k = 0; k < 3; ++k) a[k] *= s;
      return a;
    }
    friend Vec operator*(Vec a, double s) { return s * a; }
  };

  inline double gdot(Vec const& a, Vec const& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
  inline double gnorm_sq(Vec const& a) { return gdot(a, a); }
  inline double gnorm(Vec const& a) { return std::sqrt(gnorm_sq(a)); }

  struct Mat {   // 3x3, m(r, c); columns are basis vectors when used as a box basis
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    static Mat Zero() { return {}; }
    double& operator()(int r, int c) { return m[r][c]; }
    double operator()(int r, int c) const { return m[r][c]; }
  };

  template <typename T>
  struct Arr {
    std::array<T, 3> a{};
    Arr() = default;
    Arr(T x, T y, T z) : a{x, y, z} {}
    static Arr Constant(T c) { return {c, c, c}; }
    T& operator[](int i) { return a[static_cast<std::size_t>(i)]; }
    T operator[](int i) const { return a[static_cast<std::size_t>(i)]; }
  };

  // ---- properties (property.hpp:53-312) ------------------------------------------------------------
  namespace system {
    template <typename Scalar, int Rows = 1>
    struct Property {
      using scalar_t = Scalar;
      static constexpr int size() { return Rows; }
    };
  }  // namespace system

  // Each property carries its GSD chunk label (property.hpp:103-230), used by io::BinaryFile.
  struct Position : system::Property<double, 3> {
    static constexpr char const* tag = "particles/position";
  };
  struct Axis : system::Property<double, 3> {
    static constexpr char const* tag = "log/particles/axis";
  };
  struct PotentialGradient : system::Property<double, 3> {
    static constexpr char const* tag = "log/particles/potential_gradient";
  };
  struct Delta : system::Property<double, 3> {
    static constexpr char const* tag = "log/particles/delta";
  };
  struct TypeID : system::Property<std::uint32_t> {
    static constexpr char const* tag = "particles/typeid";
  };
  struct Frozen : system::Property<bool> {
    static constexpr char const* tag = "log/particles/frozen";
  };
  struct Index : system::Property<std::int64_t> {
    static constexpr char const* tag = "log/particles/index";
  };
  struct Mass : system::Property<double> {
    static constexpr char const* tag = "particles/mass";
  };
  // A species name: 32 characters, NUL padded (property.hpp:103-105).
  struct Type : system::Property<char, 32> {
    static constexpr char const* tag = "particles/types";
  };

  inline constexpr Position r_{};
  inline constexpr Axis ax_{};
  inline constexpr PotentialGradient g_{};
  inline constexpr Delta del_{};
  inline constexpr TypeID id_{};
  inline constexpr Frozen fzn_{};

  namespace system {

    // ---- SoA (SoA.hpp:74-347): one contiguous array per property, 3-vectors interleaved per atom.
    // Owning (SoA<Position>) and reference (SoA<Position&>, SoA<Position const&>) flavours share their
    // storage when converted, as the reference's views alias the caller's arrays.
    namespace detail {
      template <typename P>
      using base_t = std::remove_cv_t<std::remove_reference_t<P>>;

      template <typename Tag>
      struct Slot {
        using elem = std::conditional_t<std::is_same_v<typename Tag::scalar_t, bool>, std::uint8_t, typename Tag::scalar_t>;
        std::shared_ptr<std::vector<elem>> data = std::make_shared<std::vector<elem>>();
      };

      template <typename Tag, typename... Ps>
      inline constexpr bool has = (std::is_same_v<Tag, base_t<Ps>> || ...);
    }  // namespace detail

    // A 3-vector stored inside an SoA array.
    struct VecRef {
      double* p;
      operator Vec() const { return {p[0], p[1], p[2]}; }
      VecRef& operator=(Vec const& v) {
        p[0] = v[0];
        p[1] = v[1];
        p[2] = v[2];
        return *this;
      }
      VecRef& operator+=(Vec const& v) {
        for (int k = 0; k < 3; ++k) p[k] += v[k];
        return *this;
      }
      VecRef& operator-=(Vec const& v) {
        for (int k = 0; k < 3; ++k) p[k] -= v[k];
        return *this;
      }
      double& operator[](int i) const { return p[i]; }
    };

    // Flat view of one property array (the reference's array_ref_t).
    template <typename T>
    struct FlatView {
      T* p;
      std::size_t n;
      T* data() const { return p; }
      std::size_t size() const { return n; }
      T& operator[](std::size_t i) const { return p[i]; }
    };

    template <typename... Ps>
    class SoA : private detail::Slot<detail::base_t<Ps>>... {
      template <typename... Q>
      friend class SoA;

      template <typename Tag>
      detail::Slot<Tag>& slot() {
        return static_cast<detail::Slot<Tag>&>(*this);
      }
      template <typename Tag>
      detail::Slot<Tag> const& slot() const {
        return static_cast<detail::Slot<Tag> const&>(*this);
      }

      std::int64_t m_size = 0;

    public:
      explicit SoA(std::int64_t n = 0) { destructive_resize(n); }

      // View / copy-free conversion from any SoA that holds every property of this one.
      template <typename... Q, typename = std::enable_if_t<(detail::has<detail::base_t<Ps>, Q...> && ...)>>
      SoA(SoA<Q...> const& other) : m_size(other.size()) {
        ((slot<detail::base_t<Ps>>().data = other.template slot<detail::base_t<Ps>>().data), ...);
      }

      std::int64_t size() const { return m_size; }

      void destructive_resize(std::int64_t n) {
        m_size = n;
        ((slot<detail::base_t<Ps>>().data->resize(static_cast<std::size_t>(n * detail::base_t<Ps>::size()))), ...);
      }

      // Point property Tag at other's storage (SoA::rebind, SoA.hpp).
      template <typename Tag, typename... Q>
      void rebind(Tag, SoA<Q...> const& other) {
        slot<Tag>().data = other.template slot<Tag>().data;
      }

      template <typename Tag>
      decltype(auto) operator()(Tag, std::int64_t i) const {
        auto* base = slot<Tag>().data->data();
        if constexpr (Tag::size() == 3) {
          return VecRef{base + 3 * i};
        } else {
          return (base[i]);
        }
      }

      template <typename Tag>
      auto operator[](Tag) const {
        auto& v = *slot<Tag>().data;
        using E = typename detail::Slot<Tag>::elem;
        return FlatView<E>{v.data(), v.size()};
      }
    };

    // ---- TypeMap (typemap.hpp:67-214): a name per species and the values of the per-species
    // properties Extra... (scalar properties only, e.g. TypeMap<Mass>).
    template <typename... Extra>
    class TypeMap {
      static_assert(((Extra::size() == 1) && ...), "TypeMap holds scalar per-species properties");

    public:
      explicit TypeMap(int n = 0) : m_names(static_cast<std::size_t>(n)), m_extra(n) {}
      int num_types() const { return static_cast<int>(m_names.size()); }

      // Name and the Extra... values of species i (typemap.hpp:160-200); names must be shorter than
      // Type::size() characters, as the reference's fixed-width name storage requires.
      void set(int i, std::string name, typename Extra::scalar_t const&... values) {
        set_name(i, std::move(name));
        ((m_extra(Extra{}, i) = values), ...);
      }
      void set_name(int i, std::string name) {
        check(i);
        if (name.size() >= static_cast<std::size_t>(Type::size())) {
          throw RuntimeError("Type name length " + std::to_string(name.size()) + " but must be < " +
                             std::to_string(Type::size()));
        }
        m_names[static_cast<std::size_t>(i)] = std::move(name);
      }
      std::string const& get(int i) const { return m_names.at(static_cast<std::size_t>(i)); }
      std::string const& get(int i, Type) const { return get(i); }
      template <typename Tag, typename = std::enable_if_t<!std::is_same_v<Tag, Type>>>
      typename Tag::scalar_t get(int i, Tag) const {
        check(i);
        return m_extra(Tag{}, i);
      }

      // The per-species values as one SoA of num_types() rows (for io::BinaryFile).
      SoA<Extra...> const& values() const { return m_extra; }
      SoA<Extra...>& values() { return m_extra; }

    private:
      void check(int i) const {
        if (i < 0 || i >= num_types()) throw RuntimeError("TypeMap: species " + std::to_string(i) + " out of range");
      }
      std::vector<std::string> m_names;
      SoA<Extra...> m_extra;
    };

    // ---- Box (box.hpp:103-120) -------------------------------------------------------------------
    class Box {
    public:
      Box(Mat const& basis, Arr<bool> const& periodic) {
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) m_basis[3 * r + c] = basis(r, c);
        for (int k = 0; k < 3; ++k) m_periodic[k] = periodic[k] ? 1 : 0;
      }
      double const* basis_data() const { return m_basis; }
      int const* periodic_data() const { return m_periodic; }
      bool periodic(int i) const { return m_periodic[i] != 0; }
      Mat basis() const {
        Mat m;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) m(r, c) = m_basis[3 * r + c];
        return m;
      }

    private:
      double m_basis[9];
      int m_periodic[3];
    };

    // ---- Hessian (hessian.hpp:42-148): dense 3N x 3N, filled column by column from HVPs ----------
    class Hessian {
    public:
      void zero_for(std::int64_t n) {
        m_n = n;
        m_h.assign(static_cast<std::size_t>(9 * n * n), 0.0);
      }
      // Block (i, j) entry (r, c).
      double operator()(std::int64_t i, std::int64_t j, int r, int c) const {
        return m_h[static_cast<std::size_t>((3 * i + r) * 3 * m_n + 3 * j + c)];
      }
      double* data() { return m_h.data(); }
      std::int64_t size() const { return m_n; }

    private:
      std::int64_t m_n = 0;
      std::vector<double> m_h;
    };

  }  // namespace system
}  // namespace fly

// fly::io::BinaryFile (GSD files) reads and writes the system types above; the Options::fout hooks of the
// minimiser and dimer below write frames through it.
#define FLY_GPU_FACADE_SYSTEM_TYPES 1
#include "gsd_file.hpp"

namespace fly {

  namespace gpu {
    // Options::fout of the minimiser and dimer (lbfgs.cpp:75-77, dimer.cpp:88-90): the positions at the
    // start of every iteration committed as one frame; an exception from the file stops the run and is
    // re-thrown by the call.
    struct FrameDump {
      io::BinaryFile* file;
      std::exception_ptr error = nullptr;

      static int call(void* user, int, double const* xyz, std::int64_t n) {
        auto* d = static_cast<FrameDump*>(user);
        try {
          system::SoA<Position> s(n);
          std::copy(xyz, xyz + 3 * n, s[r_].data());
          d->file->commit([&] { d->file->write(r_, s); });
          return 0;
        } catch (...) {
          d->error = std::current_exception();
          return 1;
        }
      }
    };
  }  // namespace gpu

  // ---- neigh::List (list.hpp:71-210) -------------------------------------------------------------
  namespace neigh {
    class List {
    public:
      List(system::Box const& box, double r_cut) : m_ctx(gpu::context()), m_r_cut(r_cut) {
        gpu::check(flyg_list_create(m_ctx, box.basis_data(), box.periodic_data(), r_cut, &m_h), m_ctx);
      }
      List(List const&) = delete;
      List& operator=(List const&) = delete;
      List(List&& o) noexcept { swap(o); }
      List& operator=(List&& o) noexcept {
        swap(o);
        return *this;
      }
      ~List() {
        if (m_h) flyg_list_destroy(m_h);
      }

      auto rebuild(system::SoA<Position const&> positions, int = 1) -> void {
        m_n = positions.size();
        gpu::check(flyg_list_rebuild(m_h, positions[r_].data(), m_n, FLYG_HOST), m_ctx);
        m_cache_rc = -1;
      }

      auto update(system::SoA<Delta const&> x) -> void {
        if (x.size() != m_n) throw RuntimeError("Input to update has the wrong number of atoms");
        gpu::check(flyg_list_update(m_h, x[del_].data(), m_n, FLYG_HOST), m_ctx);
        m_cache_rc = -1;
      }

      // Calls f(n, r, dr) for every neighbour n of atom i within r_cut (list.hpp:138-154).
      template <typename F>
      auto for_neighbours(std::int64_t i, double r_cut, F&& f) const -> void {
        if (r_cut > m_r_cut) throw RuntimeError("r_cut is bigger than the internal r_cut");
        refresh(r_cut);
        for (std::int64_t k = m_off[static_cast<std::size_t>(i)]; k < m_off[static_cast<std::size_t>(i) + 1]; ++k) {
          Vec dr{m_dr[3 * k], m_dr[3 * k + 1], m_dr[3 * k + 2]};
          f(static_cast<std::int64_t>(m_j[static_cast<std::size_t>(k)]), std::sqrt(gnorm_sq(dr)), dr);
        }
      }

      auto size() const noexcept -> std::int64_t { return m_n; }
      double r_cut() const { return m_r_cut; }
      flyg_list handle() const { return m_h; }

    private:
      flyg_ctx m_ctx = nullptr;
      flyg_list m_h = nullptr;
      double m_r_cut = 0;
      std::int64_t m_n = 0;
      mutable double m_cache_rc = -1;
      mutable std::vector<std::int64_t> m_off;
      mutable std::vector<std::int32_t> m_j, m_s;
      mutable std::vector<double> m_dr;

      void swap(List& o) {
        std::swap(m_ctx, o.m_ctx);
        std::swap(m_h, o.m_h);
        std::swap(m_r_cut, o.m_r_cut);
        std::swap(m_n, o.m_n);
        std::swap(m_cache_rc, o.m_cache_rc);
        m_off.swap(o.m_off);
        m_j.swap(o.m_j);
        m_s.swap(o.m_s);
        m_dr.swap(o.m_dr);
      }

      void refresh(double rc) const {
        if (m_cache_rc == rc) return;
        std::int64_t np = 0;
        gpu::check(flyg_list_count(m_h, rc, &np), m_ctx);
        m_off.resize(static_cast<std::size_t>(m_n + 1));
        m_j.resize(static_cast<std::size_t>(np + 1));
        m_s.resize(static_cast<std::size_t>(np + 1));
        m_dr.resize(static_cast<std::size_t>(3 * np + 3));
        gpu::check(flyg_list_export(m_h, rc, m_off.data(), m_j.data(), m_s.data(), m_dr.data()), m_ctx);
        m_cache_rc = rc;
      }
    };
  }  // namespace neigh

  // ---- potentials (data.hpp, eam.hpp, generic.hpp) ------------------------------------------------
  namespace potential {

    class DataEAM {
    public:
      struct Options {
        bool debug = false;
        bool symmetric = false;
      };

      // DataEAM(Options, std::ifstream) (data.hpp:54): the file is parsed and tabulated on the device.
      DataEAM(Options const&, std::ifstream in) : m_ctx(gpu::context()) {
        if (!in.good()) throw RuntimeError("Could not open eam in");
        std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        gpu::check(flyg_eam_load_text(m_ctx, text.data(), text.size(), &m_h), m_ctx);
        int nt = 0;
        gpu::check(flyg_eam_info(m_h, &nt, &m_rcut, nullptr, nullptr), m_ctx);
        std::vector<char> names(static_cast<std::size_t>(16 * nt));
        gpu::check(flyg_eam_info(m_h, nullptr, nullptr, nullptr, names.data()), m_ctx);
        m_map = system::TypeMap<>(nt);
        for (int i = 0; i < nt; ++i) m_map.set(i, std::string(names.data() + 16 * i));
      }
      DataEAM(DataEAM const&) = delete;
      ~DataEAM() {
        if (m_h) flyg_eam_destroy(m_h);
      }

      auto r_cut() const noexcept -> double { return m_rcut; }
      auto type_map() const noexcept -> system::TypeMap<> const& { return m_map; }
      flyg_eam handle() const { return m_h; }

    private:
      flyg_ctx m_ctx;
      flyg_eam m_h = nullptr;
      double m_rcut = 0;
      system::TypeMap<> m_map;
    };

    class EAM {
    public:
      EAM(system::TypeMap<> const& map, std::shared_ptr<DataEAM const> data) : m_data(std::move(data)) {
        if (map.num_types() < m_data->type_map().num_types()) throw RuntimeError("Not enough types in eam data file");
        for (int i = 0; i < m_data->type_map().num_types(); ++i)
          if (map.get(i) != m_data->type_map().get(i)) throw RuntimeError("TypeID mismatch");
      }

      auto r_cut() const noexcept -> double { return m_data->r_cut(); }

      auto energy(system::SoA<TypeID const&, Frozen const&> in, neigh::List const& nl, int = 1) -> double {
        double E = 0;
        gpu::check(flyg_eam_energy(m_data->handle(), nl.handle(), in[id_].data(), in[fzn_].data(), in.size(), FLYG_HOST, &E));
        return E;
      }

      auto gradient(system::SoA<PotentialGradient&> out, system::SoA<TypeID const&, Frozen const&> in,
                    neigh::List const& nl, int = 1) -> void {
        if (in.size() != out.size()) throw RuntimeError("EAM gradient size mismatch");
        gpu::check(flyg_eam_gradient(m_data->handle(), nl.handle(), in[id_].data(), in[fzn_].data(), in.size(),
                                     FLYG_HOST, out[g_].data()));
      }

      // Dense Hessian (eam.cpp:122-233), assembled column by column from device Hessian-vector products.
      auto hessian(system::Hessian& out, system::SoA<TypeID const&, Frozen const&> in, neigh::List const& nl, int = 1)
          -> void {
        std::int64_t n = in.size();
        out.zero_for(n);
        std::vector<double> v(static_cast<std::size_t>(3 * n), 0.0), hv(static_cast<std::size_t>(3 * n));
        for (std::int64_t c = 0; c < 3 * n; ++c) {
          v[static_cast<std::size_t>(c)] = 1.0;
          gpu::check(flyg_eam_hvp(m_data->handle(), nl.handle(), in[id_].data(), in[fzn_].data(), n, v.data(),
                                  hv.data(), FLYG_HOST));
          v[static_cast<std::size_t>(c)] = 0.0;
          for (std::int64_t r = 0; r < 3 * n; ++r)
            out.data()[static_cast<std::size_t>(r * 3 * n + c)] = hv[static_cast<std::size_t>(r)];
        }
      }

      DataEAM const& data() const { return *m_data; }

    private:
      std::shared_ptr<DataEAM const> m_data;
    };

    // potential::Generic (generic.hpp:50-140): the EAM alternative of the reference's variant.
    class Generic {
    public:
      explicit Generic(EAM eam) : m_eam(std::move(eam)) {}
      auto r_cut() const noexcept -> double { return m_eam.r_cut(); }
      auto energy(system::SoA<TypeID const&, Frozen const&> in, neigh::List const& nl, int threads = 1) -> double {
        return m_eam.energy(in, nl, threads);
      }
      auto gradient(system::SoA<PotentialGradient&> out, system::SoA<TypeID const&, Frozen const&> in,
                    neigh::List const& nl, int threads = 1) -> void {
        m_eam.gradient(out, in, nl, threads);
      }
      auto hessian(system::Hessian& out, system::SoA<TypeID const&, Frozen const&> in, neigh::List const& nl,
                   int threads = 1) -> void {
        m_eam.hessian(out, in, nl, threads);
      }
      flyg_eam handle() const { return m_eam.data().handle(); }

    private:
      EAM m_eam;
    };

  }  // namespace potential

  // ---- minimisers (lbfgs_core.hpp, lbfgs.hpp) -----------------------------------------------------
  namespace minimise {

    class StepLBFGS {
    public:
      explicit StepLBFGS(int n) : m_ctx(gpu::context()) { gpu::check(flyg_steplbfgs_create(m_ctx, n, &m_h), m_ctx); }
      StepLBFGS(StepLBFGS const&) = delete;
      ~StepLBFGS() {
        if (m_h) flyg_steplbfgs_destroy(m_h);
      }
      auto clear() -> void { gpu::check(flyg_steplbfgs_clear(m_h), m_ctx); }

      template <typename T1, typename T2>
      auto newton_step(system::SoA<T1 const&> x, system::SoA<T2 const&> g) -> system::SoA<Delta>& {
        auto xv = x[T1{}];
        auto gv = g[T2{}];
        if (xv.size() != gv.size()) throw RuntimeError("LBFGS core inputs size mismatch");
        m_r.destructive_resize(x.size());
        gpu::check(flyg_steplbfgs_newton_step(m_h, xv.data(), gv.data(), static_cast<std::int64_t>(xv.size()),
                                              m_r[del_].data(), FLYG_HOST),
                   m_ctx);
        return m_r;
      }

    private:
      flyg_ctx m_ctx;
      flyg_steplbfgs m_h = nullptr;
      system::SoA<Delta> m_r;
    };

    class LBFGS {
    public:
      struct Options {
        int n = 10;
        int iter_max = 2000;
        double f2norm = 1e-6;
        double skin_frac = 1.1;
        double proj_tol = 0;
        double max_trust = 0.50;
        double min_trust = 0.01;
        double grow_trust = 1.5;
        double shrink_trust = 0.5;
        bool debug = false;
        io::BinaryFile* fout = nullptr;   // positions at every iteration, one frame each (lbfgs.hpp:80-87)
      };

      LBFGS(Options const& opt, system::Box const& box) : m_ctx(gpu::context()), m_opt(opt) {
        flyg_lbfgs_options o{opt.n, opt.iter_max, opt.f2norm, opt.skin_frac, opt.proj_tol,
                             opt.max_trust, opt.min_trust, opt.grow_trust, opt.shrink_trust};
        gpu::check(flyg_lbfgs_create(m_ctx, &o, box.basis_data(), box.periodic_data(), &m_h), m_ctx);
      }
      LBFGS(LBFGS const&) = delete;
      ~LBFGS() {
        if (m_h) flyg_lbfgs_destroy(m_h);
      }

      // Returns true on FAILURE (lbfgs.cpp:114).
      auto minimise(system::SoA<Position&, PotentialGradient&> out, system::SoA<Position const&, TypeID const&, Frozen const&> in,
                    potential::Generic& pot, int = 1) -> bool {
        auto o = out[r_];
        auto i = in[r_];
        for (std::size_t k = 0; k < o.size(); ++k) o[k] = i[k];
        int failed = 0, it = 0;
        gpu::FrameDump dump{m_opt.fout};
        gpu::check(flyg_lbfgs_set_observer(m_h, m_opt.fout ? &gpu::FrameDump::call : nullptr, &dump), m_ctx);
        int const rc = flyg_lbfgs_minimise(m_h, pot.handle(), o.data(), out[g_].data(), in[id_].data(), in[fzn_].data(),
                                           in.size(), FLYG_HOST, &failed, &it);
        (void)flyg_lbfgs_set_observer(m_h, nullptr, nullptr);
        if (dump.error) std::rethrow_exception(dump.error);
        gpu::check(rc, m_ctx);
        return failed != 0;
      }

      Options const& options() const { return m_opt; }

    private:
      flyg_ctx m_ctx;
      Options m_opt;
      flyg_lbfgs m_h = nullptr;
    };
  }  // namespace minimise

  // ---- saddle (rotor.hpp, dimer.hpp) -------------------------------------------------------------
  // ---- env: self-symmetries of a local environment (env/geometry.hpp:125-281, catalogue.cpp:33-64) ------
  namespace env {

    // A symmetry of a geometry: the permutation `perm` and orthogonal map O with O x[perm[i]] ~ x[i]
    // (Catalogue::SelfSymetry, catalogue.hpp:157-160).
    struct SelfSymmetry {
      Mat O;
      std::vector<std::size_t> perm;
    };

    namespace detail {
      using P3 = std::array<double, 3>;

      inline double dist(P3 const& a, P3 const& b) {
        double const x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
        return std::sqrt(x * x + y * y + z * z);
      }

      inline bool invert(Mat const& a, Mat& inv) {
        double const det = a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) -
                           a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                           a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
        if (!(std::abs(det) > 1e-300)) return false;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            int const r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
            inv(r, c) = (a(r1, c1) * a(r2, c2) - a(r1, c2) * a(r2, c1)) / det;
          }
        return true;
      }

      // ortho_onto (geometry.hpp:125-160): the orthogonal O minimising sum |O x_i - y_i|^2, reflections
      // allowed.  It is the orthogonal polar factor of M = sum y_i x_i^T (for H = sum x_i y_i^T = U S V^T it
      // equals V U^T), found by the Newton iteration X <- (X + X^-T) / 2 from X = M (Higham, Functions of
      // Matrices, ch. 8).  False when M is singular (a planar geometry).
      inline bool ortho_onto(std::vector<P3> const& x, std::vector<P3> const& y, Mat& O) {
        Mat X;
        for (std::size_t i = 0; i < x.size(); ++i)
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) X(r, c) += y[i][static_cast<std::size_t>(r)] * x[i][static_cast<std::size_t>(c)];
        for (int it = 0; it < 100; ++it) {
          Mat inv, next;
          if (!invert(X, inv)) return false;
          double change = 0, size = 0;
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
              next(r, c) = 0.5 * (X(r, c) + inv(c, r));
              change += std::abs(next(r, c) - X(r, c));
              size += std::abs(next(r, c));
            }
          X = next;
          if (change <= 1e-15 * size) break;
        }
        O = X;
        return true;
      }

      // grmsd (geometry.hpp:100-123): sqrt(sum |O x_i - y_i|^2).
      inline double grmsd(Mat const& O, std::vector<P3> const& x, std::vector<P3> const& y) {
        double s = 0;
        for (std::size_t i = 0; i < x.size(); ++i)
          for (int r = 0; r < 3; ++r) {
            double v = -y[i][static_cast<std::size_t>(r)];
            for (int c = 0; c < 3; ++c) v += O(r, c) * x[i][static_cast<std::size_t>(c)];
            s += v * v;
          }
        return std::sqrt(s);
      }

      // GREEDY_PERM (for_equiv_perms_impl, geometry.hpp:186-229): place an atom of matching colour at
      // position n, keep it while its distances to the atoms already placed match the reference's within
      // sqrt(2) delta, recurse; a complete permutation counts when its ortho_onto map brings it within
      // delta (grmsd) of the reference, and every such permutation is collected.  Any permutation that
      // counts passes every pair's test (|d(y_i, y_j) - d(x_i, x_j)| <= r_i + r_j <= sqrt(2) grmsd), so
      // testing all placed atoms (the reference tests the first ten) finds the same set while pruning the
      // search far harder; positions are taken nearest-first for the same reason.
      struct Greedy {
        std::vector<P3> const& ref;
        std::vector<int> const& col;
        double delta;
        std::vector<std::size_t> idx;   // idx[k]: the atom placed at position k
        std::vector<P3> mut;
        std::vector<SelfSymmetry> out;
        long nodes = 0;

        bool within_tol(std::size_t n) const {
          double const tol = delta * std::sqrt(2.0);
          for (std::size_t i = 0; i < n; ++i)
            if (std::abs(dist(ref[n], ref[i]) - dist(mut[n], mut[i])) > tol) return false;
          return true;
        }
        void rec(std::size_t n) {
          if (++nodes > 50000000L) throw RuntimeError("self_symmetries: search too large (delta too loose)");
          if (n >= ref.size()) {
            Mat O;
            if (ortho_onto(mut, ref, O) && grmsd(O, mut, ref) < delta) out.push_back(SelfSymmetry{O, idx});
            return;
          }
          for (std::size_t i = n; i < ref.size(); ++i) {
            if (col[idx[i]] != col[n]) continue;
            std::swap(idx[n], idx[i]);
            std::swap(mut[n], mut[i]);
            if (within_tol(n)) rec(n + 1);
            std::swap(idx[n], idx[i]);
            std::swap(mut[n], mut[i]);
          }
        }
      };
    }  // namespace detail

    // Catalogue::calc_self_syms (catalogue.cpp:33-64): every (O, perm) mapping the geometry onto itself
    // within delta, atom 0 (the centre) fixed.  `pos` relative to the geometry's centroid, `colour` per atom.
    inline std::vector<SelfSymmetry> self_symmetries(std::vector<std::array<double, 3>> const& pos,
                                                     std::vector<int> const& colour, double delta) {
      if (pos.size() != colour.size() || pos.empty()) throw RuntimeError("self_symmetries: bad geometry");
      // Nearest-first order (atom 0 stays first); the permutations come back in the caller's order.
      std::vector<std::size_t> ord(pos.size());
      std::iota(ord.begin(), ord.end(), std::size_t{0});
      std::stable_sort(ord.begin() + 1, ord.end(), [&](std::size_t a, std::size_t b) {
        return detail::dist(pos[a], pos[0]) < detail::dist(pos[b], pos[0]);
      });
      std::vector<std::array<double, 3>> p2;
      std::vector<int> c2;
      for (std::size_t k : ord) {
        p2.push_back(pos[k]);
        c2.push_back(colour[k]);
      }
      detail::Greedy g{p2, c2, delta, {}, p2, {}};
      g.idx.resize(p2.size());
      std::iota(g.idx.begin(), g.idx.end(), std::size_t{0});
      g.rec(1);
      std::vector<SelfSymmetry> out;
      for (auto const& s : g.out) {
        SelfSymmetry t{s.O, std::vector<std::size_t>(pos.size())};
        for (std::size_t k = 0; k < ord.size(); ++k) t.perm[ord[k]] = ord[s.perm[k]];
        out.push_back(std::move(t));
      }
      return out;
    }

    // The symmetries of the local environment of `centre` (atoms within r_env, minimum image, colour
    // 2 type + frozen, Geometry::centre's centroid origin; delta = 0.4 r_min as Catalogue::calc_delta with
    // an unbounded delta_max), expressed on `atoms` (e.g. the atoms a Master::Mech runs over): each kept
    // symmetry maps every one of them onto one of them of the same colour within delta (per atom), and
    // perm[i] is the atom whose image lands on atom i.  Symmetries of the environment that the wider set
    // breaks are dropped.
    inline std::vector<SelfSymmetry> site_symmetries(system::Box const& box,
                                                     system::SoA<Position const&, TypeID const&, Frozen const&> const& in,
                                                     std::int64_t centre, std::vector<std::int64_t> const& atoms,
                                                     double r_env = 5.2) {
      auto pos = in[r_];
      auto rel = [&](std::int64_t i) {
        std::array<double, 3> d{};
        for (int c = 0; c < 3; ++c) {
          double v = pos[3 * i + c] - pos[3 * centre + c];
          double const L = box.basis_data()[4 * c];
          if (box.periodic(c)) v -= L * std::round(v / L);
          d[static_cast<std::size_t>(c)] = v;
        }
        return d;
      };
      auto colour = [&](std::int64_t i) { return 2 * static_cast<int>(in[id_][i]) + (in[fzn_][i] ? 1 : 0); };
      std::vector<std::array<double, 3>> geo{rel(centre)};
      std::vector<int> col{colour(centre)};
      for (std::int64_t i = 0; i < in.size(); ++i) {
        if (i == centre) continue;
        auto d = rel(i);
        if (detail::dist(d, {0, 0, 0}) < r_env) {
          geo.push_back(d);
          col.push_back(colour(i));
        }
      }
      std::array<double, 3> cen{};
      for (auto const& p : geo)
        for (std::size_t c = 0; c < 3; ++c) cen[c] += p[c] / static_cast<double>(geo.size());
      for (auto& p : geo)
        for (std::size_t c = 0; c < 3; ++c) p[c] -= cen[c];
      double r_min = 1e300;
      for (std::size_t i = 0; i < geo.size(); ++i)
        for (std::size_t j = i + 1; j < geo.size(); ++j) r_min = std::min(r_min, detail::dist(geo[i], geo[j]));
      double const delta = 0.4 * r_min;
      std::vector<SelfSymmetry> env_syms = self_symmetries(geo, col, delta);
      // On `atoms`: positions about the same centroid.
      std::vector<std::array<double, 3>> x;
      std::vector<int> xc;
      for (std::int64_t a : atoms) {
        auto d = rel(a);
        for (std::size_t c = 0; c < 3; ++c) d[c] -= cen[c];
        x.push_back(d);
        xc.push_back(colour(a));
      }
      std::vector<SelfSymmetry> out;
      for (auto const& s : env_syms) {
        SelfSymmetry t{s.O, std::vector<std::size_t>(x.size(), x.size())};
        bool ok = true;
        for (std::size_t a = 0; a < x.size() && ok; ++a) {
          std::array<double, 3> y{};
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) y[static_cast<std::size_t>(r)] += s.O(r, c) * x[a][static_cast<std::size_t>(c)];
          std::size_t best = x.size();
          for (std::size_t b = 0; b < x.size(); ++b)
            if (xc[b] == xc[a] && detail::dist(x[b], y) < delta) {
              best = b;
              break;
            }
          if (best == x.size() || t.perm[best] != x.size()) ok = false;
          else t.perm[best] = a;
        }
        if (ok) out.push_back(std::move(t));
      }
      return out;
    }

  }  // namespace env

  namespace saddle {

    class Rotor {
    public:
      struct Options {
        int n = 6;
        int iter_max_rot = 20;
        double delta_r = 0.001;
        double theta_tol = 1 * M_PI / 360.;
        bool relax_in_convex = true;
        bool debug = false;
      };
      explicit Rotor(Options const& opt) : m_opt(opt), m_ctx(gpu::context()) {
        flyg_rotor_options r{opt.n, opt.iter_max_rot, opt.delta_r, opt.theta_tol, opt.relax_in_convex ? 1 : 0};
        gpu::check(flyg_rotor_create(m_ctx, &r, &m_h), m_ctx);
      }
      Rotor(Rotor const&) = delete;
      ~Rotor() {
        if (m_h) flyg_rotor_destroy(m_h);
      }

      // Rotor::eff_gradient (rotor.hpp:105-111): nl ready at the positions, cut-off >= r_cut(pot); the
      // list is moved by +-delta_r and returned to the same state.  Returns the curvature along the axis.
      auto eff_gradient(system::SoA<PotentialGradient&> out, system::SoA<Axis&> inout,
                        system::SoA<TypeID const&, Frozen const&> in, potential::Generic& pot, neigh::List& nl,
                        int count_frozen, int = 1) -> double {
        if (in.size() != out.size() || in.size() != inout.size()) throw RuntimeError("Rotor size mismatch");
        double curv = 0;
        gpu::check(flyg_rotor_eff_gradient(m_h, pot.handle(), nl.handle(), in[id_].data(), in[fzn_].data(), in.size(),
                                           inout[ax_].data(), out[g_].data(), count_frozen, FLYG_HOST, &curv),
                   m_ctx);
        return curv;
      }

      auto r_cut(potential::Generic const& pot) const noexcept { return pot.r_cut() + m_opt.delta_r * 2; }
      Options const& options() const { return m_opt; }

    private:
      Options m_opt;
      flyg_ctx m_ctx;
      flyg_rotor m_h = nullptr;
    };

    class Dimer {
    public:
      struct Options {
        int n = 10;
        int max_steps = 1000;
        int hist_check_freq = 5;
        double f2norm = 1e-5;
        double skin_frac = 1.1;
        double proj_tol = 0;
        double max_trust = 0.5;
        double min_trust = 0.05;
        double grow_trust = 1.5;
        double shrink_trust = 0.5;
        int convex_max = 3;
        bool use_history = true;
        bool debug = false;
        io::BinaryFile* fout = nullptr;   // centre positions at every iteration, one frame each (dimer.hpp:90-97)
      };

      // The reference's values (dimer.hpp:123-129); the C-ABI reports them without the sign.
      enum Exit : int { success = 0, convex = -1, iter_max = -2, collision = -3, uninit = -4 };

      Dimer(Options const& opt, Rotor::Options const& ropt, system::Box const& box)
          : m_ctx(gpu::context()), m_opt(opt), m_ropt(ropt) {
        flyg_dimer_options o{opt.n, opt.max_steps, opt.hist_check_freq, opt.f2norm, opt.skin_frac, opt.proj_tol,
                             opt.max_trust, opt.min_trust, opt.grow_trust, opt.shrink_trust, opt.convex_max,
                             opt.use_history ? 1 : 0};
        flyg_rotor_options r{ropt.n, ropt.iter_max_rot, ropt.delta_r, ropt.theta_tol, ropt.relax_in_convex ? 1 : 0};
        gpu::check(flyg_dimer_create(m_ctx, &o, &r, box.basis_data(), box.periodic_data(), &m_h), m_ctx);
      }
      Dimer(Dimer const&) = delete;
      ~Dimer() {
        if (m_h) flyg_dimer_destroy(m_h);
      }

      auto find_sp(system::SoA<Position&, Axis&> out, system::SoA<Position const&, Axis const&, TypeID const&, Frozen const&> in,
                   system::SoA<Position const&> in_min, potential::Generic& pot,
                   std::vector<system::SoA<Position>> const& hist_sp, double theta_tol, int count_frozen, int = 1) -> Exit {
        std::int64_t n = in.size();
        auto ro = out[r_];
        auto ao = out[ax_];
        auto ri = in[r_];
        auto ai = in[ax_];
        for (std::size_t k = 0; k < ro.size(); ++k) {
          ro[k] = ri[k];
          ao[k] = ai[k];
        }
        std::vector<double> hist;
        for (auto const& sp : hist_sp) {
          auto v = sp[r_];
          hist.insert(hist.end(), v.data(), v.data() + v.size());
        }
        flyg_search_stats st{};
        gpu::FrameDump dump{m_opt.fout};
        gpu::check(flyg_dimer_set_observer(m_h, m_opt.fout ? &gpu::FrameDump::call : nullptr, &dump), m_ctx);
        int const rc = flyg_dimer_find_sp(m_h, pot.handle(), ro.data(), ao.data(), in[id_].data(), in[fzn_].data(), n,
                                          in_min[r_].data(), hist.empty() ? nullptr : hist.data(),
                                          static_cast<int>(hist_sp.size()), theta_tol, count_frozen, FLYG_HOST, &st);
        (void)flyg_dimer_set_observer(m_h, nullptr, nullptr);
        if (dump.error) std::rethrow_exception(dump.error);
        gpu::check(rc, m_ctx);
        m_last = st;
        return static_cast<Exit>(-st.exit_code);
      }

      flyg_search_stats const& last_stats() const { return m_last; }
      Options const& options() const { return m_opt; }
      Rotor::Options const& rotor_options() const { return m_ropt; }

    private:
      flyg_ctx m_ctx;
      Options m_opt;
      Rotor::Options m_ropt;
      flyg_dimer m_h = nullptr;
      flyg_search_stats m_last{};
    };

    // Master (find.hpp:260-306, find.cpp:178-194): the saddle-search core of Master::find_mechs, with the
    // searches of a site run as one device batch.  find_mechs itself (the catalogue's canonical geometries)
    // needs the reference's env layer and is not part of this facade; find_batch below is its per-site step
    // (find.cpp:684-755) with the mechanism checks of saddle_2_mech (find.cpp:757-837), the prefactor of the
    // mass-weighted Hessians (find.cpp:838-871, 287), is_new_mech's de-duplication and append_syms's
    // symmetric copies (with env::site_symmetries in place of the catalogue's self-symmetries).
    class Master {
    public:
      struct Options {
        double r_pert = 4;
        double stddev = 0.6;
        double basin_tol = 0.05;
        double mech_tol = 0.1;
        double stationary_tol = 0.025;   // basin_tol / 2
        double hessian_eigen_zero_tol = 1e-6;
        int num_threads = 1;
        int max_searches = 250;
        int max_failed_searches = 100;
        int batch_size = 100;            // searches of a site that run as one device batch
        double nudge_frac = 0.01;
        bool debug = false;
        // Device form: atoms farther than this from a search's centre are frozen during that search.
        double active_radius = 9.0;
        // Seed of the searches' Xoshiro256** streams; 0 seeds from std::random_device as the reference does.
        std::uint64_t seed = 0;
      };

      // A mechanism over a site's atoms (its search region, ascending index): the saddle and the forward
      // minimum relative to the reverse one, and the unit dimer axis (env::Mechanism's delta_sp, delta_fwd,
      // axis, find.cpp:816-834).
      struct Mech {
        std::vector<double> delta_sp, delta_fwd, axis;
      };

      // One search of find_batch: its exit, saddle, and (for a min->sp->min pathway) the mechanism data.
      struct Search {
        Dimer::Exit exit;
        flyg_search_stats stats;
        flyg_mechanism mech;            // mech.code 0: a pathway (saddle_2_mech's checks passed)
        double kinetic_pre = 0;         // Mechanism::kinetic_pre of a pathway (find.cpp:287), Hz
        bool is_new = false;            // a pathway not within mech_tol of an earlier one at the site
        std::vector<double> sp, axis;   // 3n saddle positions and dimer axis
      };

      // Master::is_new_mech (find.cpp:1216-1277): a mechanism is known when, for some earlier one, the
      // saddle and forward displacements are both within mech_tol (root of the summed squared per-atom
      // differences), the saddle displacements point within 10 degrees and the axes within 5 (either sign).
      bool is_new_mech(Mech const& maybe, std::vector<Mech> const& hist) const {
        return is_new_mech(maybe, hist, m_opt.mech_tol);
      }
      static bool is_new_mech(Mech const& maybe, std::vector<Mech> const& hist, double mech_tol) {
        for (Mech const& m : hist) {
          double d_sp = 0, d_fwd = 0, dot = 0, n1 = 0, n2 = 0, dot_ax = 0;
          for (std::size_t i = 0; i < m.delta_sp.size(); ++i) {
            double a = maybe.delta_sp[i] - m.delta_sp[i], b = maybe.delta_fwd[i] - m.delta_fwd[i];
            d_sp += a * a;
            d_fwd += b * b;
            n1 += maybe.delta_sp[i] * maybe.delta_sp[i];
            n2 += m.delta_sp[i] * m.delta_sp[i];
            dot += maybe.delta_sp[i] * m.delta_sp[i];
            dot_ax += maybe.axis[i] * m.axis[i];
          }
          double theta = std::acos(std::clamp(dot / (std::sqrt(n1) * std::sqrt(n2)), -1., 1.)) / (2 * M_PI) * 360;
          double ax_theta = std::acos(std::clamp(dot_ax, -1., 1.)) / (2 * M_PI) * 360;
          ax_theta = std::min(ax_theta, 180 - ax_theta);
          if (std::sqrt(d_sp) < mech_tol && std::sqrt(d_fwd) < mech_tol && theta < 10 && ax_theta < 5) return false;
        }
        return true;
      }

      // Master::append_syms (find.cpp:1279-1303): for each symmetry (O, perm) of the mechanism's atoms, the
      // mechanism with atom i taking O times atom perm[i]'s axis, saddle and forward displacements is appended
      // when is_new_mech finds it new.  Returns how many were appended.  The symmetries come from
      // env::site_symmetries over the same atoms.
      std::size_t append_syms(std::vector<env::SelfSymmetry> const& syms, Mech const& new_mech,
                              std::vector<Mech>& mechs) const {
        return append_syms(syms, new_mech, mechs, m_opt.mech_tol);
      }
      static std::size_t append_syms(std::vector<env::SelfSymmetry> const& syms, Mech const& new_mech,
                                     std::vector<Mech>& mechs, double mech_tol) {
        Mech mech = new_mech;
        std::size_t count = 0;
        for (auto const& s : syms) {
          if (s.perm.size() * 3 != new_mech.delta_sp.size()) throw RuntimeError("append_syms: symmetry and mechanism sizes differ");
          for (std::size_t i = 0; i < s.perm.size(); ++i) {
            std::size_t const j = s.perm[i];
            for (int r = 0; r < 3; ++r) {
              double a = 0, d = 0, f = 0;
              for (int c = 0; c < 3; ++c) {
                a += s.O(r, c) * new_mech.axis[3 * j + static_cast<std::size_t>(c)];
                d += s.O(r, c) * new_mech.delta_sp[3 * j + static_cast<std::size_t>(c)];
                f += s.O(r, c) * new_mech.delta_fwd[3 * j + static_cast<std::size_t>(c)];
              }
              mech.axis[3 * i + static_cast<std::size_t>(r)] = a;
              mech.delta_sp[3 * i + static_cast<std::size_t>(r)] = d;
              mech.delta_fwd[3 * i + static_cast<std::size_t>(r)] = f;
            }
          }
          if (is_new_mech(mech, mechs, mech_tol)) {
            mechs.push_back(mech);
            ++count;
          }
        }
        return count;
      }

      // The masses that weight the Hessians are the potential's (setfl species masses); `map` names types.
      Master(Options const& opt, system::Box const& box, system::TypeMap<Mass> const&, potential::Generic const& pot,
             minimise::LBFGS const& min, Dimer const& dimer)
          : m_opt(opt), m_box(box), m_pot(pot.handle()), m_ctx(gpu::context()), m_min(min.options()),
            m_dimer(dimer.options()), m_rotor(dimer.rotor_options()) {
        std::uint64_t s = opt.seed ? opt.seed : (static_cast<std::uint64_t>(std::random_device{}()) << 32) ^ std::random_device{}();
        m_seed = s;
      }

      // Master::find_batch (find.cpp:684-755) at `centre` after `tot` searches there: batch_size searches
      // from Master::perturb starts (re-perturbed up to three times after a spline error), Dimer::find_sp
      // with `cache` as the known saddles and theta_tol = ((30 - 2.5) e^(-0.02 tot) + 2.5) degrees, then
      // saddle_2_mech's two relaxations and checks for each converged search.  Converged saddles are
      // appended to `cache`, as find_batch caches them for the site's later batches.
      // With `mechs` (the site's mechanisms so far), every pathway is checked by is_new_mech and new ones
      // are appended (Search::is_new).
      auto find_batch(int tot, std::int64_t centre, system::SoA<Position const&, TypeID const&, Frozen const&> in,
                      std::vector<system::SoA<Position>>& cache, std::vector<Mech>* mechs = nullptr)
          -> std::vector<Search> {
        std::int64_t const n = in.size();
        int const k = m_opt.batch_size;
        flyg_dimer_options dopt{m_dimer.n, m_dimer.max_steps, m_dimer.hist_check_freq, m_dimer.f2norm,
                                m_dimer.skin_frac, m_dimer.proj_tol, m_dimer.max_trust, m_dimer.min_trust,
                                m_dimer.grow_trust, m_dimer.shrink_trust, m_dimer.convex_max,
                                m_dimer.use_history ? 1 : 0};
        flyg_rotor_options ropt{m_rotor.n, m_rotor.iter_max_rot, m_rotor.delta_r, m_rotor.theta_tol,
                                m_rotor.relax_in_convex ? 1 : 0};
        flyg_batch_options bopt;
        flyg_batch_default_options(&bopt);
        bopt.precision = FLYG_FP64;   // the reference's precision
        flyg_batch b = nullptr;
        gpu::check(flyg_batch_create(m_ctx, m_pot, &dopt, &ropt, &bopt, m_box.basis_data(), m_box.periodic_data(),
                                     in[r_].data(), in[id_].data(), n, &b),
                   m_ctx);
        struct Guard {
          flyg_batch b;
          ~Guard() { flyg_batch_destroy(b); }
        } guard{b};
        flyg_mech_options mo;
        flyg_mech_default_options(&mo);
        mo.enabled = 1;
        mo.nudge_frac = m_opt.nudge_frac;
        mo.basin_tol = m_opt.basin_tol;
        mo.stationary_tol = m_opt.stationary_tol;
        mo.minimiser = flyg_lbfgs_options{m_min.n, m_min.iter_max, m_min.f2norm, m_min.skin_frac, m_min.proj_tol,
                                          m_min.max_trust, m_min.min_trust, m_min.grow_trust, m_min.shrink_trust};
        gpu::check(flyg_batch_set_mechanisms(b, &mo), m_ctx);
        double theta = 0;
        gpu::check(flyg_theta_tol(tot, &theta), m_ctx);
        std::vector<double> hist;
        for (auto const& sp : cache) {
          auto v = sp[r_];
          hist.insert(hist.end(), v.data(), v.data() + v.size());
        }
        std::vector<std::int64_t> centres(static_cast<std::size_t>(k), centre), ids(static_cast<std::size_t>(k));
        std::vector<double> radii(static_cast<std::size_t>(k), m_opt.active_radius);
        std::vector<std::uint64_t> states(static_cast<std::size_t>(4 * k));
        for (auto& w : states) w = splitmix64(m_seed);   // one independent Xoshiro256** state per search
        gpu::check(flyg_batch_submit_perturbed(b, k, centres.data(), radii.data(), states.data(), m_opt.stddev,
                                               m_opt.r_pert, theta, hist.empty() ? nullptr : hist.data(),
                                               static_cast<int>(cache.size()), ids.data()),
                   m_ctx);
        gpu::check(flyg_batch_run(b), m_ctx);
        std::vector<Search> out(static_cast<std::size_t>(k));
        for (int q = 0; q < k; ++q) {
          Search& r = out[static_cast<std::size_t>(q)];
          r.sp.resize(static_cast<std::size_t>(3 * n));
          r.axis.resize(static_cast<std::size_t>(3 * n));
          gpu::check(flyg_batch_result(b, ids[q], &r.stats, r.sp.data(), r.axis.data()), m_ctx);
          r.exit = r.stats.exit_code <= 4 ? static_cast<Dimer::Exit>(-r.stats.exit_code) : Dimer::uninit;
          r.mech = flyg_mechanism{-1, 0, 0.0, 0.0, {0.0, 0.0, 0.0}};
          if (r.stats.exit_code != FLYG_EXIT_SUCCESS) continue;
          std::vector<double> rev(static_cast<std::size_t>(3 * n)), fwd(static_cast<std::size_t>(3 * n));
          gpu::check(flyg_batch_mechanism(b, ids[q], &r.mech, rev.data(), fwd.data()), m_ctx);
          system::SoA<Position> c(n);
          auto cv = c[r_];
          for (std::size_t i = 0; i < cv.size(); ++i) cv[i] = r.sp[i];
          cache.push_back(std::move(c));
          if (r.mech.code != 0) continue;
          r.kinetic_pre = prefactor(centre, in, r.sp);
          if (mechs) {
            Mech m;
            double len = 0;
            for (std::int64_t i : site_atoms(centre, in)) {
              for (int k = 0; k < 3; ++k) {
                std::size_t e = static_cast<std::size_t>(3 * i + k);
                m.delta_sp.push_back(r.sp[e] - rev[e]);
                m.delta_fwd.push_back(fwd[e] - rev[e]);
                m.axis.push_back(r.axis[e]);
                len += r.axis[e] * r.axis[e];
              }
            }
            for (double& a : m.axis) a /= std::sqrt(len);
            r.is_new = is_new_mech(m, *mechs);
            if (r.is_new) mechs->push_back(std::move(m));
          }
        }
        return out;
      }

    private:
      Options m_opt;
      system::Box m_box;
      flyg_eam m_pot;
      flyg_ctx m_ctx;
      minimise::LBFGS::Options m_min;
      Dimer::Options m_dimer;
      Rotor::Options m_rotor;
      std::uint64_t m_seed;

      // splitmix64 (Vigna), the recommended way to fill a Xoshiro256** state from one 64-bit seed.
      static std::uint64_t splitmix64(std::uint64_t& x) {
        std::uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
      }

    public:
      // The atoms of a site: those within active_radius of the centre (minimum image), ascending.  A Mech
      // runs over them, and env::site_symmetries expresses a site's symmetries on them.
      std::vector<std::int64_t> site_atoms(std::int64_t centre,
                                           system::SoA<Position const&, TypeID const&, Frozen const&> const& in) const {
        std::vector<std::int64_t> out;
        auto pos = in[r_];
        double const R2 = m_opt.active_radius * m_opt.active_radius;
        for (std::int64_t i = 0; i < in.size(); ++i) {
          double d2 = 0;
          for (int c = 0; c < 3; ++c) {
            double d = pos[3 * i + c] - pos[3 * centre + c];
            double L = m_box.basis_data()[4 * c];
            if (m_box.periodic(c)) d -= L * std::round(d / L);
            d2 += d * d;
          }
          if (!(d2 > R2)) out.push_back(i);
        }
        return out;
      }

    private:
      // Mechanism::kinetic_pre (find.cpp:838-871, 287): the spectra of the mass-weighted Hessians at the
      // initial minimum and at the saddle over the search's mobile atoms (those within active_radius).
      double prefactor(std::int64_t centre, system::SoA<Position const&, TypeID const&, Frozen const&> const& in,
                       std::vector<double> const& sp) const {
        std::int64_t const n = in.size();
        auto pos = in[r_];
        std::vector<std::uint8_t> frozen(static_cast<std::size_t>(n), 1);
        for (std::int64_t i : site_atoms(centre, in)) frozen[static_cast<std::size_t>(i)] = 0;
        double rc = 0;
        int nt = 0;
        gpu::check(flyg_eam_info(m_pot, &nt, &rc, nullptr, nullptr), m_ctx);
        flyg_list nl = nullptr;
        gpu::check(flyg_list_create(m_ctx, m_box.basis_data(), m_box.periodic_data(), rc, &nl), m_ctx);
        std::vector<double> w(static_cast<std::size_t>(3 * n));
        double lp[2] = {0, 0};
        std::int64_t nz[2], nneg[2];
        double const* states[2] = {pos.data(), sp.data()};
        for (int s = 0; s < 2; ++s) {
          std::int64_t m = 0;
          int rc1 = flyg_list_rebuild(nl, states[s], n, FLYG_HOST);
          if (!rc1) rc1 = flyg_eam_hessian_eigenvalues(m_pot, nl, in[id_].data(), frozen.data(), n, FLYG_HOST, w.data(), &m);
          if (!rc1) rc1 = flyg_log_prod_eigen(w.data(), m, m_opt.hessian_eigen_zero_tol, &lp[s], &nz[s], &nneg[s]);
          if (rc1) {
            flyg_list_destroy(nl);
            gpu::check(rc1, m_ctx);
          }
        }
        flyg_list_destroy(nl);
        if (nneg[0] != 0 || nneg[1] != 1) return 0.0;   // not a minimum or not a first-order saddle
        double a = 0;
        gpu::check(flyg_prefactor(lp[0], lp[1], &a), m_ctx);
        return a;
      }
    };

  }  // namespace saddle
}  // namespace fly
