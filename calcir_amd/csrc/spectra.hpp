// spectra.hpp -- host launchers of the eigensolver, spectra and propagation
// kernels (k_eigh.hip, k_spectra.hip).
#pragma once

#include "common.hpp"

namespace cir {

// Field projection + site records and the Hamiltonian with an explicit map
// coefficient set (k_sites.hip); the cir_eproj_sites / cir_hamiltonian entry
// points resolve the set from the map table.
// boxk 0: box_d holds 3 floats (rectangular diagonal) per frame; 1: 9 floats
// (triclinic a, b, c rows).  rules: 0 for the reference's H choice and charged
// sites, or kRuleOhPerMolecule / kRuleMsiteBisector (site_rules()).
enum { kRuleOhPerMolecule = 1, kRuleMsiteBisector = 2 };
inline int site_rules(const cir_params& p) {
    return (p.oh_per_molecule ? kRuleOhPerMolecule : 0) | (p.msite_bisector ? kRuleMsiteBisector : 0);
}
int eproj_sites(int prec, int nframes, int natoms, int natom_mol, int nchrom_mol, int imodel, int rules,
                const MapCoef& mc, float cutoff, const float* xyz_d, const float* box_d, int boxk, void* sites_d,
                hipStream_t s);
int hamiltonian(int prec, int nframes, int n, int ld, int nchrom_mol, int ispecies, const MapCoef& mc, float avef,
                const float* box_d, int boxk, const void* sites_d, void* H_d, hipStream_t s);

size_t eigh_workspace(int prec, int batch, int n, int ld);
int eigh_batched(int prec, int batch, int n, int ld, void* A, void* V, double* w, void* work, hipStream_t s,
                 Prof* prof = nullptr);

// det (optional, deterministic sums): ipr_hist_det_bytes(batch, n) /
// spectral_density_det_bytes(batch, nomega) of scratch; the sums then run in
// a fixed order (per bin, frames in order, a frame's modes in order: its
// eigenvalues ascend, so the modes of one bin are contiguous) instead of by
// float atomics.
size_t ipr_hist_det_bytes(int batch, int n);
size_t spectral_density_det_bytes(int batch, int nomega);
int ipr_hist(int prec, int batch, int n, int ld, int nchrom_mol, float avef, float omegaStart, int omegaStep,
             int nomega, const void* V, const double* w, const double* weight, double* pw, double* rw,
             double* rmw, long long* skipped, hipStream_t s, void* det = nullptr);
int spectral_density(int prec, int batch, int n, int ld, float t1, float avef, float omegaStart, int omegaStep,
                     int nomega, const void* V, const double* w, const void* sites, const double* weight,
                     double* sw, double* m2work, hipStream_t s, void* det = nullptr);
// Yout = V exp(+i diag(w) dt/hbar) V^T Yin, Y row-major n x ldy with Re in
// columns [0, ncols) and Im in [ncols, 2 ncols); G: propagate_work_elems()
// reals of work.  Yout may alias Yin.  may_tune: the first call of a shape may
// time the candidate K splits (HIP events); false takes the cached or default
// split and creates nothing.
size_t propagate_work_elems(int n, int ncols);
int propagate(int prec, int gemm_prec, int n, int ld, int ncols, int ldy, double dt, const void* V,
              const double* w, const void* Yin, void* Yout, void* G, hipStream_t s, bool may_tune = true);
// Whether propagate() of this shape would time candidates (false: its split is
// settled, or the shape has none to choose), so that calls of it can be
// captured into a graph.
bool propagate_needs_tuning(int prec, int gemm_prec, int n, int ncols);
int set_identity(int prec, int n, int ldf, void* F, hipStream_t s);
int slot_fixup(int prec, int n, int nslot, int ldy, const int* action, const void* sites, const void* Yold,
               void* Ynew, hipStream_t s);
// slot_fixup then correlate in one launch (one block per slot), for frames
// with no re-unitarisation between the two.
int slot_correlate(int prec, int n, int nslot, int ldy, const int* action, const int* target, const void* sites,
                   const void* Yold, void* Ynew, double* out, hipStream_t s);
int correlate(int prec, int n, int nslot, int ldy, const int* action, const int* target, const void* sites,
              const void* Y, double* out, hipStream_t s);
int spectrum_c2r(int nspec, int nfft, int nx, const double* X, double* y, hipStream_t s);
// Re-unitarisation of the origins' propagated columns (mode per slot: 0 none,
// 1 record G0 = Y^H Y, 2 restore Y^H Y to G0); work: reunit_work_doubles().
size_t reunit_work_doubles(int nslot);
int reunitarise(int prec, int n, int nslot, int ldy, const int* mode, double* G0, void* Y, double* work,
                hipStream_t s);

}  // namespace cir
