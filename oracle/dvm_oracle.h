/* Test infrastructure only: the oracle.
 *
 * A CPU restatement of the DVMpp reference time-step path (C++17 + OpenMP,
 * FP64) used by the tests, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  Nothing in the product (dvmpp_amd/,
 * libdvm_hip.so) links or calls it.
 *
 * Every function follows a reference file:line given beside its definition in
 * dvm_oracle.cpp.  The reference itself cannot be built in this image
 * (Armadillo, Boost.program_options and BackgroundMesh.{hpp,cpp} are absent),
 * so the restatement is pinned by the reference's own known-answer tests,
 * closed forms, invariants and long-double brute force (tests/test_oracle_*.py).
 *
 * Every function has such a check.  What only the reference binary could fix
 * is the last bits of a few results, which the tests compare to a tolerance
 * instead: the summation order of arma::sum (totalcirc, the force sums), the
 * BLAS order of A^T A and A^T b and the factorisation arma::solve uses.  The
 * w component of the sheet term of orc_sheet_probe_quirk (probe_velocities'
 * altered py, qy coefficients, src/VortexSheet.cpp:611-637) is pinned by
 * rebuilding it from the panel's geometry and the quadrature-pinned
 * vortexsheetbc pair (tests/test_oracle_sheet.py).
 */
#ifndef DVM_ORACLE_H
#define DVM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Simulation parameters: the XML option map of src/XmlHandler.cpp:162-202. */
typedef struct orc_params {
    double rho;          /* constants/density */
    double nu;           /* constants/nu */
    double dt;           /* time/dt */
    double ux, uz;       /* flow/ux, flow/uz */
    double cutoff_exp;   /* constants/cutoff_exp */
    int max_npv;         /* constants/max_NumPanelVort */
    int scheme;          /* 0 = euler, 1 = RK3 (time/scheme) */
    int crossing;        /* 0 = DELETE, 1 = ABSORB, 2 = REFLECT */
    int seed;            /* constants/seed (<0: random_device) */
    /* remesh block (src/XmlHandler.cpp:192-201); frequency 0 = disabled */
    int remesh_frequency;
    int remesh_kernel;   /* 0 = M4prime, 1 = gaussian, 2 = M6 */
    double remesh_xmin, remesh_xmax, remesh_zmin, remesh_zmax;
    double remesh_dx, remesh_dz, remesh_sigma, remesh_circ_threshold;
    /* algorithms/merge_threshold, merge_frequency (src/DVMBase.cpp:49-60); 0 = off */
    double merge_threshold;
    int merge_frequency;
    /* random stream: 0 = the reference's (mt19937, src/Random.cpp:3-18);
     * 1 = the library's device stream (Philox4x32-10 keyed by seed, step and
     * blob index), so batched and device-stream runs can be checked */
    int rng_mode;
} orc_params;

/* ---- Random (src/Random.cpp:3-18) ---- */
void*    orc_rng_create(int seed);
void     orc_rng_destroy(void* rng);
double   orc_rng_rand(void* rng);
void     orc_rng_fill(void* rng, int64_t n, double* out);
uint32_t orc_rng_raw(void* rng);
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): out = the 4 words for
 * counter ctr[4] and key key[2]; orc_u01 = the 53-bit uniform in (0, 1) the
 * device stream builds from two words. */
void     orc_philox4x32(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
double   orc_u01(uint32_t a, uint32_t b);
/* OpenMP threads the parallel loops use (OMP_NUM_THREADS). */
int      orc_num_threads(void);

/* ---- Biot-Savart direct sum (src/VelocityKernel.cpp:30-108) ---- */
void orc_bs_self(int64_t n, const double* x, const double* z, const double* circ,
                 const double* sigma, double* u, double* w);
void orc_bs_targets(int64_t ns, const double* sx, const double* sz, const double* scirc,
                    const double* ssigma, int64_t nt, const double* tx, const double* tz,
                    double* u, double* w);
/* Same sums in long double, used only to pin the FP64 restatement. */
void orc_bs_self_ld(int64_t n, const double* x, const double* z, const double* circ,
                    const double* sigma, double* u, double* w);

/* ---- Vortex sheet (src/VortexSheet.cpp) ---- */
void* orc_sheet_create(int npts, const double* px, const double* pz, const orc_params* p);
void  orc_sheet_destroy(void* sh);
int   orc_sheet_size(void* sh);
/* xc zc theta ds enx enz etx etz, each Nl long; aref scalar */
void  orc_sheet_geometry(void* sh, double* xc, double* zc, double* theta, double* ds,
                         double* enx, double* enz, double* etx, double* etz, double* aref);
void  orc_sheet_influence(void* sh, double* A /* (Nl+1) x Nl, row-major */);
void  orc_sheet_get_gamma(void* sh, double* gamma);
void  orc_sheet_set_gamma(void* sh, const double* gamma);
/* solvevortexsheet: RHS b (Nl+1) and gamma */
void  orc_sheet_solve(void* sh, int64_t n, const double* bx, const double* bz,
                      const double* bcirc, const double* bsigma, double* b_out, double* gamma_out);
void  orc_sheet_bc(void* sh, int64_t n, const double* bx, const double* bz,
                   double* uvs, double* wvs);
int   orc_sheet_inside(void* sh, double x, double z);
void  orc_sheet_loads(void* sh, double ur, double* fx, double* fz, double* cd, double* cl);
/* probe_velocities as written (src/VortexSheet.cpp:559-646), quirks included */
void  orc_sheet_probe_quirk(void* sh, int64_t n, const double* bx, const double* bz,
                            const double* bcirc, const double* bsigma, int64_t np,
                            const double* px, const double* pz, double* u, double* w,
                            double* uvs, double* wvs);

/* ---- Full simulation (src/DVMBase.cpp) ---- */
void*   orc_sim_create(int npts, const double* px, const double* pz, const orc_params* p);
void    orc_sim_destroy(void* sim);
int64_t orc_sim_size(void* sim);
int     orc_sim_step(void* sim);   /* 0 ok, 3 = VortexException (NaN release) */
/* fields: id (uint32) then x z circ sigma u w uvs wvs (double) */
void    orc_sim_get_state(void* sim, uint32_t* id, double* x, double* z, double* circ,
                          double* sigma, double* u, double* w, double* uvs, double* wvs);
void    orc_sim_set_state(void* sim, int64_t n, const uint32_t* id, const double* x,
                          const double* z, const double* circ, const double* sigma,
                          const double* u, const double* w, const double* uvs,
                          const double* wvs);
void    orc_sim_get_gamma(void* sim, double* gamma);
void    orc_sim_set_gamma(void* sim, const double* gamma);
void    orc_sim_get_loads(void* sim, double* fx, double* fz, double* cd, double* cl);
int     orc_sim_step_index(void* sim);
uint64_t orc_sim_rng_draws(void* sim);
/* Puts the simulation at step `step` with the random stream `draws` doubles
 * past the seed (the stream re-seeded, then 2 words per draw discarded), so a
 * state taken from another run at that step continues as that run would. */
void    orc_sim_set_clock(void* sim, int step, uint64_t draws);
/* Individual sub-steps on the sim's state, in compute_step order. */
void    orc_sim_convect(void* sim);
void    orc_sim_remesh(void* sim);
void    orc_sim_diffuse(void* sim);
int     orc_sim_release(void* sim);
void    orc_sim_crossing(void* sim);
void    orc_sim_loads(void* sim);
/* merge_blobs on the sim's state when due (src/DVMBase.cpp:126-132); returns merges */
int64_t orc_sim_merge(void* sim);

/* ---- VortexBlobs::merge_blobs (src/VortexBlobs.cpp:110-209) ---- */
/* In place on n blobs: survivors compacted to the front, IDs renumbered
 * 0..m-1 when anything merged; returns the number of merges (n - m). */
int64_t orc_merge_blobs(double threshold, int64_t n, uint32_t* id, double* x, double* z,
                        double* circ, double* sigma, double* u, double* w, double* uvs,
                        double* wvs);

/* ---- Remesh (BackgroundMesh spec; reference source missing) ---- */
/* Returns the output count; out arrays must hold the returned size (call with
 * NULL outputs first to size them). */
int64_t orc_remesh(const orc_params* p, int64_t n, const double* x, const double* z,
                   const double* circ, const double* sigma, double* ox, double* oz,
                   double* ocirc, double* osigma);
double  orc_remesh_weight(int kernel, double r);

#ifdef __cplusplus
}
#endif
#endif
