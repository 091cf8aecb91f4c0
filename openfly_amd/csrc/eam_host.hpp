// Internal handle types of libflygpu (host side).
#pragma once

#include <istream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace flyg {

  // Raw setfl contents (host, FP64).
  struct SetflData {
    int ntypes = 0;
    std::vector<std::string> names;
    std::vector<double> mass;
    double drho = 0, dr = 0, rcut = 0;
    std::vector<std::vector<double>> F, phi, V;   // phi[a + n*b], V[max + n*min]
  };

  SetflData read_setfl(std::string const& path);
  SetflData read_setfl_stream(std::istream& in);
  std::vector<double4> spline_coefficients(std::vector<double> y, double dx);

}  // namespace flyg

// ------------------------------------------------------------------------------------------------
// Handles (the opaque types of flygpu.h).
// ------------------------------------------------------------------------------------------------

struct flyg_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cus = 0;
  flyg::DBuf<unsigned> dflags;   // [0] error bits, [1..] scratch counters
  flyg::DBuf<double> dscal;      // scalar scratch for reductions
  flyg::DBuf<double> dpart;      // block partial sums
  unsigned* hflags = nullptr;    // pinned mirror of small results
  double* hscal = nullptr;

  int refs = 1;                  // the context lives until its creator and every handle released it

  void set_device() const;
  unsigned take_flags();         // sync, read and clear the error bits
  void retain() { ++refs; }
  void release();
};

struct flyg_eam_s {
  flyg_ctx_s* ctx = nullptr;
  flyg::SetflData data;
  flyg::EAMTables tab{};   // device view
  flyg::DBuf<double4> c64;
  flyg::DBuf<float4> c32;
  int precision = FLYG_FP64;
  // scratch
  flyg::DBuf<double4> posf;     // positions + F' (pass 1 -> pass 2)
  flyg::DBuf<double2> dfp;      // HVP: (F', dF') per atom
  flyg::DBuf<double> eatom;     // per-atom energies
  flyg::DBuf<uint32_t> types;   // staged host inputs
  flyg::DBuf<uint32_t> gtype;   // types in the list's slot order (gradients with several species)
  long types_stamp = 0;         // renewed (process-wide unique) whenever the staged host types change
  long gtype_build = -1, gtype_stamp = -1;   // gtype holds the staged types in the layout of this build
  flyg::DBuf<uint8_t> frozen;
  flyg::DBuf<double> vec_in, vec_out;
  std::vector<uint32_t> last_types;
  std::vector<uint8_t> last_frozen;
  int lanes = 0;                                   // lanes per atom row, 0 = timed choice
  int staged_single = -1;                          // one species in the staged host types, else -1
  long staged_n = -1;                              // atoms in the staged copy (types/frozen), -1 = none
  std::map<std::pair<long, int>, int> tuned;       // (atoms, precision) -> lanes: density | force << 8
  // Optional kernel timing of gradient calls: one event pair around each call's kernels.
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  std::size_t events_used = 0;
  ~flyg_eam_s() {
    for (auto& e : events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
};

struct flyg_list_s {
  flyg_ctx_s* ctx = nullptr;
  flyg::BoxDesc box{};
  flyg::GridDesc grid{};
  double rc = 0;
  long n = 0;
  long nghost = 0;
  int K = 0;       // row capacity of the neighbour table
  bool built = false;
  long build_id = 0;   // unique per successful rebuild (process-wide), names the slot layout in use
  flyg::DBuf<double4> pos;       // live positions (x, y, z, 0)
  flyg::DBuf<double4> ent;       // reals + ghost images at build time (chained positions)
  flyg::DBuf<unsigned> eid;      // packed (shift, real) of ent
  flyg::DBuf<int> ecell;         // cell of ent
  flyg::DBuf<int> ccount, cstart, ccursor;
  flyg::DBuf<double4> spos;      // ent sorted by cell (packed-id order inside a cell)
  flyg::DBuf<unsigned> usid;     // packed ids scattered by cell, unordered inside a cell
  flyg::DBuf<int> rslot;         // sorted slot of each real atom
  flyg::DBuf<unsigned> sid;
  flyg::DBuf<unsigned> nbr;      // n x K packed entries
  flyg::DBuf<int> ncnt;          // entries per atom
  // The gradient passes' layout, made at each rebuild: the real atoms in cell order (slot k holds atom
  // order[k], rank[i] is atom i's slot), their positions by slot (kept current by update), and the rows
  // by slot with neighbour indices given as slots (the same entries in the same order as nbr).
  flyg::DBuf<int> order, rank;
  // gpos's w slot holds the species of each slot for a multi-species gradient (written with the slot-order
  // types; update keeps it): those of the staged host types with stamp wtype_stamp, if wtype_build is the
  // current build_id, else of the last gradient call's types.
  flyg::DBuf<double4> gpos;
  long wtype_stamp = -1, wtype_build = -1;
  flyg::DBuf<unsigned> grow;
  flyg::DBuf<int> gcnt;
  flyg::DBuf<int> scan_tmp;
  flyg::DBuf<double> stage;      // host staging of xyz vectors
  flyg::DBuf<unsigned> lstate;   // device build counters
  unsigned* hstate = nullptr;    // pinned mirror
  long gcap = 0;                 // ghost slots
  ~flyg_list_s() {
    if (hstate) (void)hipHostFree(hstate);
  }
};

namespace flyg {
  // Shift vectors of the 27 image codes (sum of +-basis columns), for pair arithmetic.
  void shift_table(BoxDesc const& b, double3* out27);

  // neigh.hip
  void list_rebuild(flyg_list_s& nl, double const* d_xyz, long n);
  void list_update(flyg_list_s& nl, double const* d_delta, long n);
  void exclusive_scan(flyg_ctx_s& ctx, int* data, int n, flyg::DBuf<int>& tmp);

  // eam.hip
  struct PairArgs {
    double4 const* pos;
    unsigned const* nbr;
    int const* cnt;
    int K;
    long n;
    uint32_t const* types;
    uint8_t const* frozen;
    int t0;   // the one species present (-1: several or unknown); read by the single-species kernels
    // Null: rows, positions and types are by atom.  Otherwise they are by slot of the list's cell-order
    // layout (types gathered into slot order) and order[slot] is the atom, for frozen flags and gradients.
    int const* order = nullptr;
    // The w slot of pos holds each atom's species (as a double), so a neighbour's species arrives with
    // its position; set for the slot layout with several species.
    bool type_in_w = false;
  };
  void eam_setup_tables(flyg_eam_s& e);
  double eam_energy(flyg_eam_s& e, flyg_list_s& nl, uint32_t const* d_types, uint8_t const* d_frozen);
  void eam_gradient(flyg_eam_s& e, flyg_list_s& nl, uint32_t const* d_types, uint8_t const* d_frozen, double* d_grad,
                    int precision);
  // Allocation-free gradient: the pass-1 output lives in a file-scope device array of NOALLOC_MAX_ATOMS
  // entries; noalloc_scratch_init(device) looks its address up once per device (that device current).
  constexpr long NOALLOC_MAX_ATOMS = 131072;
  void noalloc_scratch_init(int device);
  void eam_gradient_noalloc(flyg_eam_s& e, flyg_list_s& nl, uint32_t const* d_types, uint8_t const* d_frozen,
                            double* d_grad, int precision);
  // Sorted eigenvalues of the mass-weighted Hessian block of the non-frozen atoms (hessian.hip).
  std::vector<double> hessian_eigenvalues(flyg_eam_s& e, flyg_list_s& nl, uint32_t const* d_types,
                                          uint8_t const* d_frozen, std::vector<uint8_t> const& h_frozen,
                                          std::vector<uint32_t> const& h_types, int* sweeps);
  void eam_hvp(flyg_eam_s& e, flyg_list_s& nl, uint32_t const* d_types, uint8_t const* d_frozen, double const* d_v,
               double* d_hv);
  void eam_table_eval(flyg_eam_s& e, int kind, int a, int b, double const* x, long m, int deriv, double* out);

  // Stage host per-atom inputs into the potential handle (re-uploads only when they changed); device
  // inputs pass through; types = frozen = null selects the staged copy.
  void stage_atoms(flyg_eam_s& e, uint32_t const* types, uint8_t const* frozen, long n, int where,
                   uint32_t const** d_types, uint8_t const** d_frozen);
  // Copy device per-atom inputs into the handle's staged buffers.
  void stage_atoms_device(flyg_eam_s& e, uint32_t const* types, uint8_t const* frozen, long n);

  // Deterministic device reductions (vecops.hip).
  void dev_sum(flyg_ctx_s& ctx, double const* x, long n, double* d_out);
}  // namespace flyg
