/*
 * calcir.h -- C ABI of libcalcir, the MI355X (gfx950) spectroscopy hot path
 * of calcIR (nicholashestand/calcIR): electric-field projection, Skinner
 * maps, exciton Hamiltonian, batched symmetric eigensolver, propagation,
 * IR/Raman correlation functions and lineshapes.
 *
 * Conventions
 *   - Every entry point returns an int status (cir_status); none exits the
 *     process.  cir_strerror() names a status; cir_last_error() returns the
 *     detail text of the most recent failure on the calling thread.  No
 *     C++ exception crosses the ABI: a failed host allocation returns
 *     CIR_ERR_NOMEM.  Corrupt XTC input (bad counts or ranges in a frame)
 *     returns CIR_ERR_IO.
 *   - Pointers named *_d are device pointers (hipMalloc'd, on the current
 *     HIP device); "stream" is a hipStream_t passed as void* (NULL = the
 *     null stream).  Per-stage calls enqueue work and return without
 *     synchronising.
 *   - A cir_ctx owns its device buffers, is bound to one stream and is not
 *     thread-safe; separate contexts may run concurrently.
 *   - Real arrays of the per-stage calls hold "real_t" = float for
 *     CIR_PREC_FP32/CIR_PREC_BF16X3 and double for CIR_PREC_FP64.
 *
 * Reference boundary replaced (file:line in calcIR.cu / calcIR.h):
 *   calcIR.exe <input>                calcIR.cu:9-18     -> calcir CLI (cir_main)
 *   ir_init                            calcIR.cu:1539-1643 -> cir_params_read
 *   defaults                           calcIR.cu:36-59   -> cir_params_default
 *   model/species resolution           calcIR.cu:324-339 -> cir_resolve_model/_species
 *   sample x frame loop + finalize     calcIR.cu:360-919 -> cir_create/cir_run/cir_get_results
 *                                      (or cir_stream_begin/_push/_end with frames fed in order)
 *   get_eproj_GPU                      calcIR.cu:1020-1168, calcIR.h:51-53 -> cir_eproj_sites
 *   get_kappa_GPU                      calcIR.cu:1176-1406, calcIR.h:56-60 -> cir_eproj_sites + cir_hamiltonian
 *   magma_ssyevd_gpu                   calcIR.cu:420-448 -> cir_eigh_batched
 *   get_spectral_density + 3 sgemv     calcIR.cu:458-482, 1414-1444, calcIR.h:62-64 -> cir_spectral_density
 *   IPR / Pw,Rw,Rmw host loop          calcIR.cu:494-538 -> cir_ipr_hist
 *   cast_to_complex/makeI/Pinit + cgemm x3 + ccopy   calcIR.cu:549-617, 1481-1535,
 *                                      calcIR.h:66-76    -> cir_propagate (+ cir_gemm)
 *   15 x (cgemv + cdotu)               calcIR.cu:628-745 -> cir_correlate
 *   cufftExecC2R x 3                   calcIR.cu:775-827 -> cir_spectrum_c2r
 *   output writers                     calcIR.cu:843-919 -> cir_write_outputs
 *   minImage/mag3/dot3                 calcIR.cu:1457-1476, calcIR.h:79-88 -> cir_minimage
 *   xdrfile read_xtc* / xdr_seek       calcIR.cu:212-219, 341-395 -> cir_xtc_scan/_read/_times,
 *                                      cir_xtc_open/_next/_seek/_close, cir_xtc_offsets (+ _write)
 */
#ifndef CALCIR_H
#define CALCIR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cir_status {
    CIR_OK = 0,
    CIR_ERR_ARG = 1,        /* invalid argument or size                        */
    CIR_ERR_IO = 2,         /* input file / output file could not be used       */
    CIR_ERR_MODEL = 3,      /* model string not recognised (calcIR.cu:327-330)  */
    CIR_ERR_SPECIES = 4,    /* species not recognised (calcIR.cu:336-339)       */
    CIR_ERR_MAP = 5,        /* map is neither 0 nor 1 (calcIR.cu:1207-1242)     */
    CIR_ERR_FRAME = 6,      /* frame missing or time check failed (calcIR.cu:379-395) */
    CIR_ERR_HIP = 7,        /* HIP runtime error                                */
    CIR_ERR_NOMEM = 8,      /* device or host allocation failed                 */
    CIR_ERR_EIGH = 9,       /* eigensolver failure (calcIR.cu:445)              */
    CIR_ERR_NONFINITE = 10, /* non-finite coordinate or box                     */
    CIR_ERR_NODEVICE = 11,  /* no HIP device available                          */
    CIR_INCOMPLETE = 12     /* stopped at max_frames after writing a checkpoint */
} cir_status;

enum { CIR_PREC_FP32 = 0, CIR_PREC_FP64 = 1, CIR_PREC_BF16X3 = 2 };
enum { CIR_MODE_PLAIN = 0, CIR_MODE_BATCHED = 1 };

#define CIR_STRLEN 256
#define CIR_NSITE 20 /* real_t fields per chromophore in a site record */

/* Site-record field offsets (each field is an array of N values). */
enum {
    CIR_SITE_E = 0, CIR_SITE_W = 1, CIR_SITE_X = 2, CIR_SITE_P = 3, CIR_SITE_MUP = 4,
    CIR_SITE_U = 5,   /* 3 fields: unit OH vector                       */
    CIR_SITE_R = 8,   /* 3 fields: min-imaged TDM position (nm)         */
    CIR_SITE_MU = 11, /* 3 fields: transition dipole                    */
    CIR_SITE_A = 14   /* 6 fields: alpha xx, yy, zz, xy, yz, zx         */
};

/* Input keys of calcIR.cu:1556-1638 plus extensions. */
typedef struct cir_params {
    char  xtcf[CIR_STRLEN];    /* trajectory; "synth:<nwater>:<nframes>:<seed>[:<L_nm>]" selects the
                                  seeded synthetic source                            */
    char  outf[CIR_STRLEN];
    char  cptf[CIR_STRLEN];    /* checkpoint prefix (the reference parses it and never
                                  uses it); see ckpt_every below                     */
    char  model[CIR_STRLEN];
    char  species[CIR_STRLEN];
    int   ntcfpoints;
    int   nsamples;
    float sampleEvery;
    int   omegaStep;
    int   natom_mol;
    int   nchrom_mol;
    int   nzeros;
    int   imap;                /* key "map"                                        */
    float dt;
    float t1;
    float avef;
    float beginTime;
    float omegaStart;
    float omegaStop;
    /* extensions */
    float cutoff;              /* key "cutoff" (nm), default 0.7831 (calcIR.cu:1035) */
    int   precision;           /* key "precision": fp32 | fp64 | bf16x3 (default fp32) */
    int   mode;                /* key "mode": batched (default) | plain             */
    int   chunk;               /* key "chunk": frames per eigensolver batch, 0 = auto */
    /* checkpoint/resume (batched mode), file <cptf>_ckpt.bin: */
    int   ckpt_every;          /* key "checkpoint_every": write after at least this many
                                  distinct frames (at chunk ends), 0 = never           */
    int   resume;              /* key "resume": 1 = continue from the checkpoint        */
    int   max_frames;          /* key "max_frames": stop after this many distinct frames
                                  in this run, checkpoint and return CIR_INCOMPLETE; 0 = all */
    /* lineshape adjoint (cir_forward_ls / cir_backward_ls): */
    int   adj_store;           /* key "adjoint_store": 0 = keep every frame's eigensystem
                                  when it fits in device memory, else recompute them chunk
                                  by chunk in the reverse sweep; 1 = always keep; 2 = always
                                  recompute                                            */
    int   triclinic;           /* key "triclinic": 1 = the box arrays given to cir_run,
                                  cir_run_device and cir_forward_ls hold 9 floats per frame,
                                  the rows a, b, c of a GROMACS box (a along x, b in the xy
                                  plane, reduced); 0 = 3 floats, the rectangular diagonal as
                                  the reference (calcIR.cu:404-405). Set by the CLI when the
                                  trajectory's boxes have off-diagonal elements          */
    int   reunit_every;        /* key "reunitarise_every" (batched mode): every this many
                                  propagated points of a sample, restore the Gram matrix of
                                  its nine propagated columns, Y^H Y, to its value at the
                                  origin (F unitary keeps it there; rounding lets it drift);
                                  0 = never                                             */
    int   origin_stride;       /* key "origin_stride_frames": > 0 starts sample s this many
                                  trajectory frames after sample s-1 (sampleEvery =
                                  origin_stride * frame spacing); 0 = sampleEvery as given */
    int   oh_per_molecule;     /* key "oh_per_molecule": 1 = chromophore k of a molecule is
                                  the O-H(1+k) stretch (nchrom_mol 1 or 2), so with
                                  nchrom_mol 1 every molecule contributes its H1 stretch;
                                  0 (default) = the reference's rule, H1 or H2 by chrom % 2
                                  whatever nchrom_mol is (calcIR.cu:1061-1070, 1266-1275,
                                  1360-1370): with nchrom_mol 1, even molecules use H1 and
                                  odd molecules H2                                      */
    int   msite_bisector;      /* key "msite_bisector": 1 = with natom_mol 3, a -1.04 charge
                                  on the HOH bisector 0.015 nm from O (the TIP4P M site)
                                  enters the field; 0 (default) = the reference's charged
                                  sites, atoms 1..natom_mol-1 with +0.52 on atoms 1, 2 and
                                  -1.04 on atom 3 (calcIR.cu:1117-1160), so three-site
                                  water carries only its two H charges                  */
    float time_prec;           /* key "time_precision": the frame spacing is the first two
                                  frames' time difference rounded at this precision,
                                  round(dt * prec) / prec (calcIR.cu:342-345, where prec is
                                  the XTC file's own coordinate precision).  Default 1000
                                  (GROMACS' default); the CLI sets it from the file
                                  (cir_xtc_precision)                                   */
    int   deterministic;       /* key "deterministic": 1 = results bitwise reproducible
                                  run to run (no float atomics, no timed launch shapes:
                                  two-stage eigensolver, one K segment in the
                                  propagation, histograms and densities summed in a
                                  fixed order); needs precision fp32 or bf16x3 and
                                  66 <= N <= 2080.  Default 0                           */
} cir_params;

/* Results of a run, all arrays caller-allocated (host).  NULL members are
 * skipped. */
typedef struct cir_results {
    double* tcf_ir;   /* 2*ntcfpoints: decayed, 1/nsamples-scaled TCF (re, im) pairs */
    double* tcf_vv;
    double* tcf_vh;
    double* ftcf_ir;  /* ntcfpoints+nzeros: C2R of the padded TCF              */
    double* ftcf_vv;
    double* ftcf_vh;
    double* sw;       /* nomega each                                            */
    double* pw;
    double* rw;
    double* rmw;
    int64_t nsteps;        /* (sample, tcf point) steps                        */
    int64_t nframes_used;  /* distinct trajectory frames diagonalised          */
    int64_t hist_skipped;  /* modes outside the omega window (not binned)      */
    double  seconds;       /* device time of the last cir_run                  */
} cir_results;

typedef struct cir_ctx cir_ctx;

/* Device-time phases reported by cir_phase_times (HIP events on the
 * context's stream around each phase's launches). */
enum {
    CIR_PH_SITES = 0,         /* field projection + site records            */
    CIR_PH_HAMILTONIAN = 1,   /* exciton Hamiltonian                        */
    CIR_PH_TRIDIAG = 2,       /* Householder reduction (to tridiagonal form,
                                 or to band form where the bulge-chasing
                                 stage follows)                              */
    CIR_PH_BISECT = 3,        /* tridiagonal eigenvalues                    */
    CIR_PH_INVIT = 4,         /* tridiagonal eigenvectors                   */
    CIR_PH_BACKTRANSFORM = 5, /* eigenvectors V = Q Z (Householder panels)  */
    CIR_PH_IPR = 6,           /* IPR / frequency distribution               */
    CIR_PH_SW = 7,            /* spectral density                           */
    CIR_PH_PROPAGATE = 8,     /* propagation products + phase               */
    CIR_PH_CORRELATE = 9,     /* origin bookkeeping + correlation functions */
    CIR_PH_FINALIZE = 10,     /* C2R lineshapes                             */
    CIR_PH_BULGE = 11,        /* band -> tridiagonal by bulge chasing       */
    CIR_PH_BULGE_BT = 12,     /* bulge-chasing reflectors applied to Z      */
    CIR_NPHASE = 13
};

/* --- host-side boundary (no GPU needed) --------------------------------- */
const char* cir_strerror(int status);
const char* cir_last_error(void);
void  cir_params_default(cir_params* p);                        /* calcIR.cu:36-59 */
int   cir_params_read(const char* path, cir_params* p, int verbose);  /* calcIR.cu:1539-1643 */
void  cir_params_echo(const cir_params* p);                     /* calcIR.cu:70-87 */
int   cir_resolve_model(const char* model, int* imodel);        /* calcIR.cu:325-330 */
int   cir_resolve_species(const char* species, int* ispecies);  /* calcIR.cu:332-339 */
int   cir_nomega(const cir_params* p);                          /* calcIR.cu:94 */
float cir_minimage(float dx, float boxl);                       /* calcIR.cu:1457-1460 */
/* Frame selected for (sample, point): calcIR.cu:378-379. */
int   cir_frame_index(const cir_params* p, int sample, int point, float frame_dt, float* desired_time);
/* Writes the 13 output files of calcIR.cu:843-919 from r. */
int   cir_write_outputs(const cir_params* p, const cir_results* r);
/* Synthetic trajectories (replacing the XTC reader, calcIR.cu:212-227, 341-398). */
float cir_synth_box_nm(int nwater);
int   cir_synth_trajectory(int nwater, int nframes, uint64_t seed, float box_nm, float* xyz,
                           float* times);
/* GROMACS XTC trajectories, replacing xdrfile's read_xtc_natoms /
 * read_xtc_n_frames / read_xtc / xdr_seek (calcIR.cu:212-219, 341-395,
 * calcIR.h:16-18).  cir_xtc_scan counts atoms and frames.  cir_xtc_read
 * decodes frames first..first+count-1 into caller arrays: xyz count*natoms*3
 * nm, box count*3 (the diagonal, as calcIR.cu:404-405 uses it), times count
 * ps, steps count (box, times, steps may be NULL).  cir_xtc_write writes (or
 * with append != 0 appends) frames with the given precision (1000 = 0.001 nm,
 * GROMACS' default) and a diagonal box; frame k gets step k. */
int   cir_xtc_scan(const char* path, int* natoms, int* nframes);
int   cir_xtc_read(const char* path, int first, int count, int natoms, float* xyz, float* box, float* times,
                   int* steps);
int   cir_xtc_write(const char* path, int natoms, int nframes, const float* xyz, const float* box,
                    const float* times, float precision, int append);
/* The same with the full box matrix, 9 floats per frame (rows a, b, c; the
 * layout of the triclinic key's box arrays). */
int   cir_xtc_read_full(const char* path, int first, int count, int natoms, float* xyz, float* box9,
                        float* times, int* steps);
int   cir_xtc_write_full(const char* path, int natoms, int nframes, const float* xyz, const float* box9,
                         const float* times, float precision, int append);
/* Times and full boxes of the first nframes frames without decoding any
 * coordinates (each compressed block is skipped by its byte count): times
 * nframes ps, box9 nframes*9 (either may be NULL). */
int   cir_xtc_times(const char* path, int nframes, float* times, float* box9);
/* Sequential reading, the role of xdrfile's open + read_xtc loop
 * (calcIR.cu:341-395): cir_xtc_open starts at frame 0; cir_xtc_next decodes
 * the next count frames (xyz count*natoms*3 nm, box9 count*9, times count;
 * box9 and times may be NULL) and fails with CIR_ERR_FRAME past the end. */
typedef struct cir_xtc cir_xtc;
int   cir_xtc_open(const char* path, int natoms, cir_xtc** out);
int   cir_xtc_next(cir_xtc* reader, int count, float* xyz, float* box9, float* times);
void  cir_xtc_close(cir_xtc* reader);
/* Random access, the role of the xtc_seek extension's frame offsets and
 * xdr_seek (calcIR.cu:380): cir_xtc_offsets gives the byte offset of each of
 * the first nframes frames (from the headers and the coordinate blocks' byte
 * counts, nothing decoded); cir_xtc_seek positions a reader at a frame (the
 * index is built on its first call), after which cir_xtc_next reads from
 * there; CIR_ERR_FRAME past the last frame. */
int   cir_xtc_offsets(const char* path, int nframes, int64_t* offsets);
/* The coordinate precision stored in frame `frame` (the prec that xdrfile's
 * read_xtc returns, calcIR.cu:342-345); -1 for frames of at most 9 atoms,
 * which the format stores uncompressed without a precision. */
int   cir_xtc_precision(const char* path, int frame, float* precision);
int   cir_xtc_seek(cir_xtc* reader, int frame);
/* The drop-in CLI: argv as calcIR.exe (calcIR.cu:9-18). */
int   cir_main(int argc, char** argv);

/* --- whole pipeline ------------------------------------------------------ */
int   cir_device_count(int* n);
int   cir_create(const cir_params* p, int natoms, void* stream, cir_ctx** out);
/* Runs every sample of p over an in-memory trajectory (host arrays:
 * xyz nframes*natoms*3 nm, box nframes*3, times nframes ps). */
int   cir_run(cir_ctx* ctx, int nframes, const float* xyz, const float* box, const float* times);
/* Same with the trajectory already on the device. */
int   cir_run_device(cir_ctx* ctx, int nframes, const float* xyz_d, const float* box_d,
                     const float* times);
/* Streaming form of cir_run for trajectories that are never resident, as the
 * reference reads each frame on demand (xdr_seek + read_xtc,
 * calcIR.cu:341-395).  cir_stream_begin takes the times of all nframes
 * frames, which fix the frames every (sample, point) uses (calcIR.cu:360-395);
 * cir_stream_push hands over frames first..first+count-1 (host arrays, xyz
 * count*natoms*3 nm, box count*3, or count*9 with the triclinic key) in
 * trajectory order, and runs each chunk of used frames once its last frame
 * has arrived; cir_stream_end finishes the run (results as after cir_run)
 * and fails with CIR_ERR_FRAME if a used frame never arrived.  Batched mode
 * only, without checkpoint_every / resume / max_frames.  A failing push
 * closes the stream. */
int   cir_stream_begin(cir_ctx* ctx, int nframes, const float* times);
int   cir_stream_push(cir_ctx* ctx, int first, int count, const float* xyz, const float* box);
int   cir_stream_end(cir_ctx* ctx);
/* Independent jobs at once on one GPU (for example boxes of uneven size):
 * job j is cir_run(ctxs[j], nframes[j], xyz[j], box[j], times[j]); each job
 * runs on its own stream (a context bound to the null stream gets one for the
 * call), issued by nthreads host threads (0 = 8), largest jobs first.  The
 * jobs' device buffers come from a stream-ordered memory pool of the library
 * (hipFree would wait for every other job's queued work); the pool gives the
 * memory back before the call returns, after waiting for the jobs' streams
 * (a caller's stream bound to a context included).  The contexts must be
 * distinct.
 * Returns the first failing job's status. */
int   cir_run_many(int njob, cir_ctx* const* ctxs, const int* nframes, const float* const* xyz,
                   const float* const* box, const float* const* times, int nthreads);
int   cir_get_results(cir_ctx* ctx, cir_results* r);
/* Enables (on != 0) per-phase device timing of subsequent runs. */
int   cir_set_profiling(cir_ctx* ctx, int on);
/* Device seconds and timed launch groups per phase of the last run
 * (arrays of nphase <= CIR_NPHASE entries). */
int   cir_phase_times(cir_ctx* ctx, double* seconds, long long* groups, int nphase);

/* The launch configurations chosen so far in this process by timing on first
 * use (no counterpart in the reference: calcIR launches fixed shapes), as
 * text, one line per size bucket: family, bucket, choice, candidates' times.
 * Writes at most buflen bytes (NUL-terminated) and returns the length of the
 * whole report.  CIR_NO_TUNE=1 in the environment turns the timing off. */
int   cir_tuning_report(char* buf, int buflen);
void  cir_destroy(cir_ctx* ctx);

/* --- map coefficients and the lineshape adjoint ---------------------------- */
/* The 13 map coefficients of a (map, species) row of calcIR.cu:1203-1242, in
 * the order of CIR_COEF_*: omega = w0 + w1 E + w2 E^2, x10 = x0 + x1 omega,
 * p10 = p0 + p1 omega, mu' = mu0 + mu1 E + mu2 E^2, intramolecular coupling
 * (wi0 + wi1 (E_i + E_j)) x_i x_j + wi2 p_i p_j (calcIR.cu:1253-1341). */
#define CIR_NCOEF 13
enum {
    CIR_COEF_W0 = 0, CIR_COEF_W1, CIR_COEF_W2, CIR_COEF_X0, CIR_COEF_X1, CIR_COEF_P0, CIR_COEF_P1,
    CIR_COEF_MU0, CIR_COEF_MU1, CIR_COEF_MU2, CIR_COEF_WI0, CIR_COEF_WI1, CIR_COEF_WI2
};
int   cir_map_coefs(int imap, int ispecies, double* coef /* CIR_NCOEF */);
/* Coefficients a context uses (the table row of its map and species unless
 * overridden); cir_set_coefs(ctx, NULL) restores the table row.  Applies to
 * cir_run and cir_forward_ls. */
int   cir_get_coefs(const cir_ctx* ctx, double* coef);
int   cir_set_coefs(cir_ctx* ctx, const double* coef);
/* Forward pass of the IR lineshape: the C2R output of the decayed, averaged
 * IR correlation function (ftcf_ir, ntcfpoints+nzeros doubles; the same as
 * cir_run's ftcf_ir), keeping every distinct frame's eigensystem and the
 * propagated dipole vectors on the device for cir_backward_ls.  Host arrays
 * as cir_run.  No reference counterpart: the reference has no derivatives. */
int   cir_forward_ls(cir_ctx* ctx, int nframes, const float* xyz, const float* box, const float* times,
                     double* ftcf_ir);
/* Reverse pass: given dL/dftcf_ir (ntcfpoints+nzeros doubles) for any loss L
 * of the last forward pass's output, writes dL/dcoef (CIR_NCOEF doubles). */
int   cir_backward_ls(cir_ctx* ctx, const double* dL_dftcf_ir, double* dL_dcoef);

/* --- per-stage device entry points --------------------------------------- */
/* Field projection (calcIR.cu:1020-1168) fused with the per-row site record
 * of get_kappa_GPU (calcIR.cu:1245-1304).  xyz_d nframes*natoms*3 float,
 * box_d nframes*3 float; sites_d nframes*CIR_NSITE*N real_t. */
int cir_eproj_sites(int prec, int nframes, int natoms, int natom_mol, int nchrom_mol, int imodel,
                    int ispecies, int imap, float cutoff, const float* xyz_d, const float* box_d,
                    void* sites_d, void* stream);
/* Exciton Hamiltonian (calcIR.cu:1306-1404) from site records: H_d
 * nframes*N*ld real_t, symmetric (both triangles written), ld >= N. */
int cir_hamiltonian(int prec, int nframes, int n, int ld, int nchrom_mol, int ispecies, int imap,
                    float avef, const float* box_d, const void* sites_d, void* H_d, void* stream);
/* Batched symmetric eigensolver replacing magma_ssyevd_gpu (calcIR.cu:443):
 * A_d holds batch matrices (n x n, leading dimension ld, batch stride n*ld,
 * symmetric, both triangles) and is destroyed (it receives the Householder
 * reflectors); V_d receives the eigenvectors row-major, V[b][i*ld + k] =
 * component i of eigenvector k, and w_d (batch*n doubles) the eigenvalues in
 * ascending order.  Real arrays are float unless prec == CIR_PREC_FP64;
 * CIR_PREC_BF16X3 keeps float arrays and forms the back-transform products
 * with split-BF16x3 MFMA (eigenvector error ~1e-5 instead of ~1e-6).
 * ld must be a multiple of 4 and A_d 16-byte aligned (rows are streamed as
 * 16-byte vectors).  work_d must hold cir_eigh_workspace() bytes. */
size_t cir_eigh_workspace(int prec, int batch, int n, int ld);
int cir_eigh_batched(int prec, int batch, int n, int ld, void* A_d, void* V_d, double* w_d, void* work_d,
                     void* stream);
/* IPR / frequency distribution (calcIR.cu:494-538): for each matrix b adds
 * weight[b]/omegaStep per mode (weight_d NULL = 1) into pw/rw/rmw (nomega
 * doubles each); skipped_d (one int64) counts modes outside the window,
 * which the reference writes out of bounds (calcIR.cu:530-537). */
int cir_ipr_hist(int prec, int batch, int n, int ld, int nchrom_mol, float avef, float omegaStart,
                 int omegaStep, int nomega, const void* V_d, const double* w_d,
                 const double* weight_d, double* pw_d, double* rw_d, double* rmw_d,
                 long long* skipped_d, void* stream);
/* Spectral density (calcIR.cu:458-482, 1414-1444): sw_d += weight[b] *
 * sum_k |V^T mu|_k^2 Lorentz(omega - w_k - avef), mu from the site records;
 * work_d batch*n doubles. */
int cir_spectral_density(int prec, int batch, int n, int ld, float t1, float avef,
                         float omegaStart, int omegaStep, int nomega, const void* V_d,
                         const double* w_d, const void* sites_d, const double* weight_d,
                         double* sw_d, double* work_d, void* stream);
/* One propagation step (calcIR.cu:597-617, Pinit 1498-1516):
 * Yout = V exp(+i diag(w) dt/hbar) V^T Yin for ncols complex columns stored
 * row-major n x ldy, real parts in columns [0, ncols), imaginary parts in
 * [ncols, 2 ncols).  With ncols = n and Yin = Yout = [Re F | Im F] this is the
 * literal F-matrix step; with the 9 columns F M(0) of each time origin it is
 * the vector form.  prec CIR_PREC_BF16X3 runs the two products as split
 * bf16 x 3.  Up to 512 complex columns (float arrays) the products run as
 * one 32 x 32 output block per workgroup over all of K with the row phase in
 * the first product's epilogue; wider or FP64 products use the tiled GEMM,
 * split along K when the column count is small.  work_d must hold
 * cir_propagate_workspace() bytes. */
size_t cir_propagate_workspace(int prec, int n, int ncols);
int cir_propagate(int prec, int n, int ld, int ncols, int ldy, float dt, const void* V_d,
                  const double* w_d, const void* Yin_d, void* Yout_d, void* work_d, void* stream);
/* cir_propagate without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * allocates nothing: its work buffer is a fixed device array of
 * cir_propagate_noalloc_max_bytes() bytes (32 MiB) in the library, one per
 * GPU, held until the process ends; shapes whose cir_propagate_workspace()
 * exceeds it return CIR_ERR_ARG.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (concurrent streams or threads)
 * race.  The K split of the products is the one a cir_propagate call of the
 * same shape tuned earlier in the process, else a fixed default (no timing
 * runs, no HIP events).  Results equal cir_propagate's to rounding. */
size_t cir_propagate_noalloc_max_bytes(void);
int cir_propagate_noalloc(int prec, int n, int ld, int ncols, int ldy, float dt, const void* V_d,
                          const double* w_d, const void* Yin_d, void* Yout_d, void* stream);
/* Correlations (calcIR.cu:628-745): for each of nslot time origins s with
 * action_d[s] in {1, 2} (NULL = all), the 15 bilinear forms M_b(t)^T Y_a of
 * Y = F M(0) (slot s: columns 9s..9s+8 of Y_d as in cir_propagate) with
 * M(t) from sites_d, combined into IR, VV, VH and stored as 6 doubles at
 * out_d + 6*target_d[s] (target_d NULL = 0). */
int cir_correlate(int prec, int n, int nslot, int ldy, const int* action_d, const int* target_d,
                  const void* sites_d, const void* Y_d, double* out_d, void* stream);
/* cuFFT C2R contract (calcIR.cu:790-791) for nspec spectra: X_d nspec*nx
 * complex doubles (re, im), y_d nspec*nfft doubles. */
int cir_spectrum_c2r(int nspec, int nfft, int nx, const double* X_d, double* y_d, void* stream);
/* Re-unitarisation of propagated origins (no reference counterpart: the
 * reference's F-matrix is never corrected).  Y_d as cir_propagate's vector
 * form (n rows of [Re | Im], 9 complex columns per slot, ldy >= 18 nslot);
 * mode_d[nslot]: 0 leave the slot, 1 record G0_d[s] = Y_s^H Y_s (real, 81
 * doubles per slot; Y_s = M(0) at an origin), 2 set Y_s <- Y_s (I - G0^{-1}
 * (Y_s^H Y_s - G0)/2), which restores Y_s^H Y_s to G0 to second order. */
size_t cir_reunit_workspace(int nslot);
int cir_reunitarise(int prec, int n, int nslot, int ldy, const int* mode_d, double* G0_d, void* Y_d,
                    void* work_d, void* stream);
/* Batched dense product C = alpha A B + beta C (the library GEMMs of
 * calcIR.cu:462-467, 601-613): element (m,k) of A at A[m*a_m + k*a_k] (one of
 * the two strides 1), B (k,n) at B[k*b_k + n*b_n], C (m,n) at C[m*c_m + n*c_n],
 * batch strides a_b, b_b, c_b.  prec: FP32 MFMA, BF16X3 MFMA or FP64. */
int cir_gemm(int prec, int M, int N, int K, int batch, const void* A, long long a_m, long long a_k,
             long long a_b, const void* B, long long b_k, long long b_n, long long b_b, void* C,
             long long c_m, long long c_n, long long c_b, double alpha, double beta, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CALCIR_H */
