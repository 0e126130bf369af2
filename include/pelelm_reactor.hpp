// pelelm_reactor.hpp -- a ReactorBase-shaped C++ adapter over the C ABI in pelelm_chem.h.
//
// PeleLM holds one `std::unique_ptr<pele::physics::reactions::ReactorBase>` per process
// (Source/PeleLM.H:16,672-674) and drives it with
//   create(name)                                   Source/PeleLM.cpp:706
//   init(reactor_type, ncells)                     Source/PeleLM.cpp:707-709
//   set_typ_vals_ode(typ)                          Source/PeleLM.cpp:1513
//   react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time, stream)
//                                                  Source/PeleLM.cpp:5752-5758
//   close()                                        Source/PeleLM.cpp:930
// This header keeps those names, argument order and meanings so a PeleLM-shaped driver
// compiles against it unchanged.  AMReX is not a dependency: `mi::Box` and `mi::Array4`
// below carry the same index arithmetic as amrex::Box / amrex::Array4 (i fastest, then
// j, k, component), and a driver built with AMReX converts with `mi::view(array4)`.
//
// Errors: the C ABI returns negative codes; this adapter throws std::runtime_error with
// the reactor's message.  A cell that fails to integrate throws from react() under
// mi.on_fail = abort (the default: the reference's integrator aborts inside react, and its
// caller ignores the return value, Source/PeleLM.cpp:5752); under mi.on_fail = continue
// react() returns 0 and the failed cells carry FC < 0.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pelelm_chem.h"

namespace mi {

struct IntVect {
  int v[3];
  int operator[](int d) const { return v[d]; }
};

// Cell-centred box [lo, hi] (inclusive), like amrex::Box.
struct Box {
  IntVect lo, hi;
  IntVect smallEnd() const { return lo; }
  IntVect bigEnd() const { return hi; }
  long numPts() const {
    return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
};

// View of a device array with amrex::Array4 indexing:
// p[(i-begin.x) + (j-begin.y)*jstride + (k-begin.z)*kstride + n*nstride].
template <class T>
struct Array4 {
  T* p = nullptr;
  IntVect begin{{0, 0, 0}}, end{{0, 0, 0}};   // end is exclusive
  long jstride = 0, kstride = 0, nstride = 0;
  int ncomp = 0;
};

// An amrex::Array4-like object (members p, begin, end, jstride, kstride, nstride, ncomp)
// as the C ABI's view; T is double or int.
template <class A>
inline mi_real4 view(const A& a) {
  mi_real4 v;
  v.p = const_cast<double*>(a.p);
  for (int d = 0; d < 3; ++d) {
    v.begin[d] = a.begin.v[d];
    v.end[d] = a.end.v[d];
  }
  v.jstride = a.jstride;
  v.kstride = a.kstride;
  v.nstride = a.nstride;
  v.ncomp = a.ncomp;
  return v;
}
template <class A>
inline mi_int4 view_int(const A& a) {
  mi_int4 v;
  v.p = const_cast<int*>(a.p);
  for (int d = 0; d < 3; ++d) {
    v.begin[d] = a.begin.v[d];
    v.end[d] = a.end.v[d];
  }
  v.jstride = a.jstride;
  v.kstride = a.kstride;
  v.nstride = a.nstride;
  v.ncomp = a.ncomp;
  return v;
}

class ReactorBase {
 public:
  // ReactorBase::create -- "ReactorCvode" / "ReactorBDF" (dense analytic-Jacobian BDF on
  // the GPU), "ReactorRK64" (explicit Runge-Kutta sub-steps) or "ReactorNull"
  // (rY += dt rYsrc, rE += dt rEsrc).
  static std::unique_ptr<ReactorBase> create(const std::string& name) {
    void* h = nullptr;
    if (mi_chem_create(name.c_str(), &h) != MI_OK)
      throw std::runtime_error("unknown chemistry integrator " + name);
    return std::unique_ptr<ReactorBase>(new ReactorBase(h));
  }
  ~ReactorBase() { close(); }
  ReactorBase(const ReactorBase&) = delete;
  ReactorBase& operator=(const ReactorBase&) = delete;

  // reactor_type 2 = h-reactor (the only one PeleLM uses); ncells sizes the workspace.
  int init(int reactor_type, int ncells) { return check(mi_chem_init(h_, reactor_type, ncells)); }

  // The ParmParse keys of the reference decks (ode.rtol, ode.atol, cvode.max_order,
  // cvode.solve_type, peleLM.use_typ_vals_chem, ...) and this reactor's mi.* keys.
  void set(const std::string& key, const std::string& value) { check(mi_chem_set(h_, key.c_str(), value.c_str())); }

  void set_typ_vals_ode(const std::vector<double>& typ) {
    check(mi_chem_set_typ_vals(h_, typ.data(), (int)typ.size()));
  }

  // Advance every cell of `box` with mask >= 0 from `time` over `dt` on `stream`
  // (a hipStream_t, or nullptr for the null stream).  Returns 0; throws on any error,
  // including a failed cell under mi.on_fail = abort (the default).
  template <class R4, class I4>
  int react(const Box& box, const R4& rY, const R4& rYsrc, const R4& T, const R4& rE, const R4& rEsrc, const R4& FC,
            const I4& mask, double& dt, double& time, void* stream = nullptr) {
    int lo[3] = {box.lo[0], box.lo[1], box.lo[2]};
    int hi[3] = {box.hi[0], box.hi[1], box.hi[2]};
    const int rc = mi_chem_react_box(h_, lo, hi, view(rY), view(rYsrc), view(T), view(rE), view(rEsrc), view(FC),
                                     view_int(mask), &dt, &time, stream);
    return check(rc);
  }

  mi_chem_stats stats() {
    mi_chem_stats s;
    check(mi_chem_get_stats(h_, &s));
    return s;
  }

  int nspec() { return mi_chem_nspec(h_); }

  void close() {
    if (h_) mi_chem_close(h_);
    h_ = nullptr;
  }

 private:
  explicit ReactorBase(void* h) : h_(h) {}
  int check(int rc) {
    if (rc != MI_OK) {
      const char* m = mi_chem_last_error(h_);
      throw std::runtime_error(std::string("pelelm_chem error ") + std::to_string(rc) + ": " + (m ? m : ""));
    }
    return rc;
  }
  void* h_;
};

}  // namespace mi
