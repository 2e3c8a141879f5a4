// Device-side description of the molecule for the Fock kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace mhip {

// Atom-pair tables built once per geometry (see SystemTables in context.cpp).
struct SysDev {
  int numat = 0;
  int norbs = 0;
  long long n2elec = 0;
  const int* nf0 = nullptr;             // first orbital of each atom (0-based)
  const int* norb = nullptr;            // 1 (H) or 4 (C, N, O)
  const long long* woff = nullptr;      // start of atom i's W row (pairs j < i)
  const long long* wcum = nullptr;      // sum_{j'<j} lim(j'): pair offset inside a row / lim(i)
  const long long* w1off = nullptr;     // one-centre block of atom i
  const long long* ccol = nullptr;      // column scratch start of atom j
  const int* heavy = nullptr;           // heavy atoms, ascending
  const int* light = nullptr;           // hydrogens, ascending
  const int* nheavy_before = nullptr;   // heavy atoms with index < i
  const int* nlight_before = nullptr;   // hydrogens with index < i
  const double* w = nullptr;            // two-electron integrals, hcore order
};

enum FockInit { kInitH = 0, kInitShift = 1, kInitRandom = 2, kInitCopy = 3 };

struct FockArgs {
  SysDev sys;
  const double* h = nullptr;    // packed one-electron matrix
  const double* p = nullptr;    // packed total density
  const double* pa = nullptr;   // packed alpha density
  const double* fin = nullptr;  // packed starting F (kInitCopy)
  double* f = nullptr;          // packed Fock (output; may alias fin)
  double* jrow = nullptr;       // numat x 10 scratch
  double* jcol = nullptr;       // column scratch
  double shift = 0.0;
  int init_mode = kInitH;
};

void fock_build(hipStream_t s, const FockArgs& a);

// ---- batched forms (one launch for many molecules) --------------------------
struct BatchCnvg {
  int n;
  double* p;     // packed total density (new, extrapolated in place)
  double* pold;  // packed previous density
  double* p1;    // diagonal history (n)
};
// args: one FockArgs per molecule (device array); blk_mol/blk_atom map each
// workgroup to (molecule, atom); mode/shift per molecule override the args.
void fock_build_batch(hipStream_t s, const FockArgs* args, int nblk, const int* blk_mol, const int* blk_atom,
                      const int* active, const int* mode, const double* shift);
void helect_batch(hipStream_t s, const FockArgs* args, int nmol, int nrow_blk, const int* blk_mol,
                  const int* blk_row, const int* active, const long long* orb_off, double* partial,
                  double scale, double* ee);
void cnvg_batch(hipStream_t s, const BatchCnvg* b, int nmol, int nrow_blk, const int* blk_mol,
                const int* blk_row, const int* active, const int* niter, double* scal);
void helect(hipStream_t s, int n, const double* pa, const double* h, const double* f, double scale,
            double* partial, double* out);
void cnvg(hipStream_t s, int n, double* pnew, double* pold, double* p1, int niter, double* scal);
void axpby(hipStream_t s, long long len, double a, const double* x, double b, double* y);

}  // namespace mhip
