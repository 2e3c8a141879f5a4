// The library context (mhip_ctx): one device, one stream, the uploaded
// system and the SCF's resident buffers.  Internal to the library; the C ABI
// sees it as an opaque pointer.
#pragma once
#include <string>
#include <vector>

#include "common.h"
#include "k_eigen.h"
#include "k_fock.h"
#include "k_guess.h"
#include "k_occvirt.h"
#include "k_pseudo.h"

namespace mhip {
struct MzFock;  // MOZYME sparse Fock build state (mozyme_scf.hip)
void mz_release(MzFock* z);
}  // namespace mhip

struct mhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // system
  int numat = 0, norbs = 0, npad = 0;
  long long mpack = 0, n2elec = 0;
  mhip::DevBuf w, h, nf0, norb, woff, wcum, w1off, ccol, heavy, light, nhb, nlb, jrow, jcol;
  mhip::SysDev sys;
  std::vector<int> nf0_h, no_h;  // first orbital (0-based) and orbital count per atom
  bool have_system = false;
  // fragment start (CSR over fragments)
  mhip::DevBuf frag_ptr, frag_orb, frag_occ, frag_ooff, frag_voff, frontier;
  // SCF state
  mhip::DevBuf p, pa, f, fcopy, pold, pold2, fd, A, C, C2, T, eig, scal, partial;
  mhip::DevBuf fhist, ehist, coef;
  mhip::DevBuf split_scratch;
  // learned starting density (mhip_guess_water), valid for the uploaded system
  mhip::GuessWork gw;
  mhip::DevBuf pguess;
  bool have_guess = false;
  // MOZYME system (mhip_mozyme_setup)
  mhip::MzFock* mz = nullptr;
  mhip::JacobiPlan plan;
  mhip::JacobiPlan plan_split;  // occupied / virtual groups, for the exit after block steps
  bool have_c = false;
  double* hscal = nullptr;  // pinned host scalars
  void* hstage = nullptr;   // pinned staging for result downloads (grow-only)
  size_t hstage_bytes = 0;
  // pseudo-diagonalization workspace
  mhip::PseudoWork pseudo;
  mhip::OccVirtWork ov;
  // statistics of the last eigensolve
  int last_sweeps = 0;
  // Jacobi stopping tolerance (relative to max|diag|) inside the SCF loop;
  // the exit diagonalization always uses 1e-12.
  double tol_iter = 1.0e-10;
  int total_sweeps = 0;
};

namespace mhip {
// pair tables and buffers of a system in the hcore integral layout (scf_driver.hip)
void build_system(mhip_ctx* c, int numat, const int* nfirst, const int* nlast, int norbs, long long n2elec,
                  const double* h, const double* w);
FockArgs fock_args(mhip_ctx* c, int mode, double shift);
void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s);
void copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s);
}  // namespace mhip
