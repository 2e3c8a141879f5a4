// Device-pointer cores of the MOZYME entry points (k_mozyme.hip), for the
// resident MOZYME loop (mozyme_scf.hip): the LMO set lives in the device
// copies between mz_resident_begin and mz_resident_end, P and F are device
// arrays of the sparse storage, and the mhip_lmo_set passed to the diagg
// cores only supplies sizes and the host copies of the LMO offsets.
#pragma once
#include <hip/hip_runtime.h>

#include "mopac_hip.h"

namespace mhip {

hipStream_t mz_stream_of(mhip_ctx* ctx);
void mz_resident_begin(hipStream_t s, const mhip_lmo_set* L);
void mz_resident_end(hipStream_t s, mhip_lmo_set* L);
bool mz_tidy_dev(hipStream_t s, int numat, int norbs, bool occ, int nmo, int* nn_host, int* no_host,
                 double thresh);
double mz_check_dev(hipStream_t s, bool occ, int nmo);
void mz_eimp_blocks(hipStream_t s, int numat, const int* nijbo, const int* iorbs, int* nblk);
void mz_eimp_dev(hipStream_t s, int nblk, const double* f_d, double* p_d);
void mz_density_dev(hipStream_t s, long long mpack, int nocc, int mode, const double* partp_d, double* p_d);
double mz_helecz_dev(hipStream_t s, int numat, const double* p_d, const double* h_d, const double* f_d);
double mz_cnvgz_dev(hipStream_t s, int norbs, long long mpack, double* pn_d, double* p_d, double* p1_d,
                    double* p2_d, double* p3_d, int niter, const int* idiag_d, int three_point);
void mz_diagg1_dev(hipStream_t s, const mhip_lmo_set* L, const double* fao_d, const double* p_d, double* eigs,
                   double* eigv, int idiagg, mhip_diagg1_state* st, int fmo_dim, double* fmo, int* ifmo, int* nij);
void mz_diagg2_dev(hipStream_t s, const mhip_lmo_set* L, const double* eigs, const double* eigv, const double* fmo,
                   const int* ifmo, int nij, int idiagg, mhip_diagg2_state* st);

}  // namespace mhip
