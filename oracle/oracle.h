/*
 * oracle.h -- CPU restatement of calcIR's spectroscopy pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (calcir_amd/, the
 * calcir CLI, libcalcir.so) links, loads or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only
 * as the checker or as the timed CPU baseline.
 *
 * Every routine follows a cited region of the reference (calcIR.cu /
 * calcIR.h of nicholashestand/calcIR).  Each routine is instantiated in
 * float (ORC_F32, the reference's user_real_t) and double (ORC_F64, the
 * error reference).
 *
 * Parity pinning (what checks each routine independently of its own text):
 *   orc_minimage, orc_frame_index   closed forms incl. half-box ties
 *                       (tests/test_oracle_geometry_field.py)
 *   orc_eproj           brute-force Coulomb sum on hand-placed molecules, cutoff
 *                       in/out, own-molecule exclusion, M relocation, periodic
 *                       image invariance (tests/test_oracle_geometry_field.py)
 *   orc_sites           closed-form map polynomials, both maps x 4 species
 *   orc_hamiltonian     symmetry, HOD => diagonal, two-molecule TDC closed form,
 *                       intramolecular closed form (tests/test_oracle_maps_hamiltonian.py)
 *   orc_eigh            numpy.linalg.eigh fixtures (tests/golden/eigh_fixtures.npz),
 *                       residual and orthogonality invariants, lower-triangle use
 *   orc_c2r             numpy.fft.irfft(norm="forward") fixtures, DC/Nyquist
 *                       imaginary parts, damped exponential closed form
 *                       (tests/test_oracle_eigh_spectrum.py)
 *   orc_spectral_density  uncoupled chromophores -> Lorentzian sum closed form
 *   orc_ipr_hist        localised modes -> IPR = 1, histogram normalisation
 *   orc_run             uncoupled (HOD) closed-form TCF, literal F-matrix vs
 *                       vector propagation, C_IR(0) = sum |mu|^2, VV(0)/VH(0) =
 *                       orientational averages of alpha, FP32 vs FP64
 *                       (tests/test_oracle_pipeline.py); coefficient override
 *                       equal to the table row reproduces the table run, and
 *                       the closed-form map polynomials hold at overridden
 *                       coefficients (tests/test_oracle_maps_hamiltonian.py)
 */
#ifndef CALCIR_ORACLE_H
#define CALCIR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F32 = 0, ORC_F64 = 1 };
enum { ORC_PROP_LITERAL = 0, ORC_PROP_VECTOR = 1 };

/* Number of doubles per chromophore in the site record written by orc_sites. */
enum {
    ORC_SITE_E = 0,      /* projected field (a.u.)                       */
    ORC_SITE_W = 1,      /* omega (cm^-1)                                */
    ORC_SITE_X = 2,      /* x10                                          */
    ORC_SITE_P = 3,      /* p10                                          */
    ORC_SITE_MUP = 4,    /* mu' (dipole derivative)                      */
    ORC_SITE_U = 5,      /* unit OH vector, 3 values                     */
    ORC_SITE_R = 8,      /* TDM position (min-imaged), 3 values (nm)     */
    ORC_SITE_MU = 11,    /* transition dipole mu = u mu' x, 3 values     */
    ORC_SITE_A = 14,     /* alpha xx,yy,zz,xy,yz,zx, 6 values            */
    ORC_SITE_STRIDE = 20
};

/* Already-resolved run settings (integers for model/species/map), as the
 * reference has them after calcIR.cu:324-339. */
typedef struct orc_config {
    int   natom_mol;     /* atoms per molecule (4)                         */
    int   nchrom_mol;    /* chromophores per molecule (2)                  */
    int   imodel;        /* 0 tip4p/e3b2, 1 tip4p2005/e3b3                 */
    int   ispecies;      /* 0 H2O, 1 HOD/D2O, 2 HOD/H2O, 3 D2O             */
    int   imap;          /* 0 Gruenbaum 2013, 1 Li-Skinner 2010            */
    float cutoff;        /* O_m - H_n field cutoff (nm), reference 0.7831  */
    float avef;          /* cm^-1                                          */
    float dt;            /* ps between TCF points                          */
    float t1;            /* ps                                             */
    int   ntcfpoints;
    int   nsamples;
    int   nzeros;
    float sampleEvery;   /* ps                                             */
    float beginTime;     /* ps                                             */
    float omegaStart;
    float omegaStop;
    int   omegaStep;
    int   propagation;   /* ORC_PROP_LITERAL or ORC_PROP_VECTOR            */
    int   nthreads;      /* OpenMP threads (0 = library default)           */
    /* map coefficients in place of the calcIR.cu:1203-1242 table row when
     * use_coef != 0: w0 w1 w2, x0 x1, p0 p1, mu0 mu1 mu2, wi0 wi1 wi2 */
    int    use_coef;
    double coef[13];
    /* box arrays hold 9 floats per frame (rows a, b, c of a GROMACS box, a
     * along x, b in the xy plane) instead of the 3-float diagonal */
    int    triclinic;
    /* oh_per_molecule != 0: chromophore k of a molecule is the O-H(1+k)
     * stretch (nchrom_mol <= 2).  0 (the reference, calcIR.cu:1061-1070,
     * 1266-1275, 1360-1370): H1 or H2 by chrom % 2, so with nchrom_mol 1
     * molecules alternate H1, H2 by parity. */
    int    oh_per_molecule;
    /* msite_bisector != 0 with natom_mol 3: a -1.04 site is built on the HOH
     * bisector 0.015 nm from O.  0 (the reference, calcIR.cu:1117-1160): the
     * charged sites are atoms 1..natom_mol-1, H (+0.52) for atoms 1, 2 and M
     * (-1.04) for atom 3, so three-site water carries no negative site. */
    int    msite_bisector;
    /* precision at which the first two frames' spacing is rounded,
     * round(dt * prec) / prec (calcIR.cu:342-345; prec is the XTC file's);
     * 0 means 1000 */
    float  time_prec;
} orc_config;

/* Outputs of orc_run.  All arrays are caller-allocated. */
typedef struct orc_result {
    /* decayed, 1/nsamples-scaled TCFs (calcIR.cu:777-781), complex as
     * (re, im) pairs, ntcfpoints each */
    double *tcf_ir, *tcf_vv, *tcf_vh;
    /* C2R outputs, n = ntcfpoints + nzeros reals each */
    double *ftcf_ir, *ftcf_vv, *ftcf_vh;
    /* nomega each, normalised as calcIR.cu:831-839 */
    double *sw, *pw, *rw, *rmw;
    /* counters */
    int64_t nsteps;          /* (sample, tcf point) steps processed        */
    int64_t hist_skipped;    /* eigenvalues outside the histogram window  */
    /* seconds per phase: field+H, eigh, spectral density + ipr, propagate
     * + correlate, finalize */
    double t_hamiltonian, t_eigh, t_hist, t_propagate, t_finalize;
} orc_result;

/* calcIR.cu:1457-1460 */
double orc_minimage(int prec, double dx, double boxl);

/* Number of omega bins, calcIR.cu:94. */
int orc_nomega(const orc_config* cfg);

/* Frame index selected for (sample, point), calcIR.cu:378-379; -1 if the
 * time check of calcIR.cu:391 would fail given frame_dt. */
int orc_frame_index(const orc_config* cfg, int sample, int point, float frame_dt,
                    float* desired_time);

/* calcIR.cu:1020-1168.  x: natoms*3 floats (nm).  eproj: nchrom doubles. */
int orc_eproj(int prec, const orc_config* cfg, const float* x, int natoms,
              const float box[3], double* eproj);

/* calcIR.cu:1245-1304 (per-row part of get_kappa_GPU).  site: nchrom *
 * ORC_SITE_STRIDE doubles. */
int orc_sites(int prec, const orc_config* cfg, const float* x, int natoms,
              const float box[3], const double* eproj, double* site);

/* calcIR.cu:1306-1404.  H: nchrom*nchrom doubles, row-major kappa[n*N+m]. */
int orc_hamiltonian(int prec, const orc_config* cfg, const float* x, int natoms,
                    const float box[3], const double* eproj, double* H);

/* Symmetric eigensolve replacing magma_ssyevd_gpu (calcIR.cu:443): the lower
 * row-major triangle of H (n >= m) is used, as MAGMA's "Upper" on the
 * column-major view reads it.  w ascending; V row-major, V[i*n+k] =
 * component i of eigenvector k. */
int orc_eigh(int prec, int n, const double* H, double* w, double* V, int nthreads);

/* Spectral density of one frame, calcIR.cu:458-482, 1414-1444.  Adds into
 * sw (nomega). w: eigenvalues (H - avef), V as orc_eigh, mu: 3*n (x,y,z
 * blocks). */
int orc_spectral_density(int prec, const orc_config* cfg, int n, const double* w,
                         const double* V, const double* mu, double* sw);

/* Frequency distribution and IPR of one frame, calcIR.cu:494-538.  Adds
 * weight/omegaStep per mode into pw, rw, rmw (nomega); returns the number of
 * modes outside the window (skipped). */
int64_t orc_ipr_hist(int prec, const orc_config* cfg, int n, const double* w,
                     const double* V, double weight, double* pw, double* rw,
                     double* rmw);

/* cuFFT C2R contract (calcIR.cu:790-791): y[k] = sum_j Xh[j] e^{+2 pi i jk/n}
 * with the Hermitian extension of X[0..n/2]; X has nx complex values
 * ((re,im) pairs), the rest of the n/2+1 inputs are zero. */
int orc_c2r(int n, int nx, const double* X, double* y);

/* Whole pipeline with the reference control flow (calcIR.cu:360-919) on an
 * in-memory trajectory: xyz nframes*natoms*3 floats, box nframes*3, times
 * nframes (ps). */
int orc_run(int prec, const orc_config* cfg, int natoms, int nframes,
            const float* xyz, const float* box, const float* times,
            orc_result* out);

#ifdef __cplusplus
}
#endif

#endif
