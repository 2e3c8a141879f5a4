// Learned starting density for water clusters, built on the device: pair
// descriptors, the MLP forward pass and the assembly of the packed density
// (k_guess.hip).  The host keeps only the neighbour list and the checks.
#pragma once
#include "common.h"

struct mhip_mlp;

namespace mhip {

// network hooks (k_mlp.hip)
float* mlp_input(mhip_mlp* q, int batch);
const float* mlp_forward_resident(mhip_mlp* q, int batch);
int mlp_layers(const mhip_mlp* q);
int mlp_dim(const mhip_mlp* q, int l);

constexpr int kGuessNin = 96;   // pair descriptor length
constexpr int kGuessNout = 16;  // one 4 x 4 density block in the molecule frames
constexpr int kGuessTypes = 8;  // pair types (self O/H, intramolecular OH/HH, intermolecular OO/OH/HO/HH)

struct GuessHost {
  int numat, nw, npairs;
  const int* nat;      // atomic numbers (1 or 8)
  const double* xyz;   // numat x 3, Angstrom
  const int* mol;      // molecule of each atom
  const int* waters;   // nw x 3: O, H1, H2
  const int* pa;       // pair atoms, pa[m] >= pb[m]
  const int* pb;
  const int* nf0;      // first orbital of each atom (0-based)
  const int* no;       // orbital count of each atom
  const float* x_mean; // kGuessNin
  const float* x_std;
  const float* y_mean; // kGuessTypes x kGuessNout
  const float* y_std;
};

struct GuessWork {
  DevBuf xyz, nat, mol, side, waters, R, geo, atomf, pa, pb, typ, nf0, no, norm;
};

// Writes the packed starting density (mpack) into p on stream s.  x_out
// (npairs x kGuessNin, the normalized descriptors) and y_out (npairs x
// kGuessNout, the raw network outputs) are copied back when not null.
void guess_water(hipStream_t s, GuessWork& g, mhip_mlp* q, const GuessHost& in, double* p,
                 long long mpack, float* x_out, float* y_out);

}  // namespace mhip
