/*
 * lmoracle.h -- C interface of the Langmuir KMC hot-path ORACLE.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU, FP64 restatement of the
 * reference's CPU path (Random, Grid, Potential, ChargeAgent, Simulation,
 * Source/Drain agents and the World construction order).  It is used by the
 * tests, by the smoke check and by the CPU-baseline leg of the benchmark, only
 * as the checker or as the timed CPU baseline.  The product library
 * (langmuir_amd/, liblangmuir_gpu.so) never links, loads or calls it.
 *
 * All integers are int32 unless stated; every function returns 0 on success
 * and a negative value on error (the message is available from
 * orc_last_error()).
 */
#ifndef LMORACLE_H
#define LMORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Subset of SimulationParameters (src/langmuirCore/include/parameters.h:46-381)
 * that the hot path and the world construction read.  Field order is the ABI. */
typedef struct orc_params {
    int32_t simulation_type;        /* 0 = transistor, 1 = solarcell */
    int32_t grid_x, grid_y, grid_z;
    int32_t coulomb_carriers;
    int32_t defects_charge;
    int32_t hopping_range;          /* 1 or 2 */
    int32_t electrostatic_cutoff;   /* sites */
    int32_t source_metropolis;
    int32_t source_coulomb;
    int32_t balance_charges;
    int32_t recombination_range;
    uint32_t random_seed;           /* must be > 0 (0 would seed from the clock) */
    int32_t reserved0;
    double coulomb_gaussian_sigma;
    double electron_percentage, hole_percentage, seed_charges, defect_percentage;
    double trap_percentage, trap_potential, gaussian_stdev, seed_percentage;
    double voltage_left, voltage_right, exciton_binding, temperature_kelvin;
    double source_rate, e_source_l_rate, e_source_r_rate, h_source_l_rate,
           h_source_r_rate, generation_rate;
    double drain_rate, e_drain_l_rate, e_drain_r_rate, h_drain_l_rate,
           h_drain_r_rate;
    double boltzmann_constant, dielectric_constant, elementary_charge,
           permittivity_space, grid_factor;
    double slope_z, recombination_rate, source_scale_area;
} orc_params;

/* flags of the stateless pair sums */
#define ORC_GAUSS 1  /* multiply each term by erf(r / (sqrt(2) sigma)) */
#define ORC_IMAGE 2  /* left-electrode image: dx = |xi + xj| + 1, term subtracted */

typedef struct orc_world orc_world;

const char* orc_last_error(void);

/* ---- engine ---------------------------------------------------------------- */
/* MT19937 seeded like boost::mt19937(seed); writes n tempered outputs. */
int orc_mt_outputs(uint32_t seed, uint64_t skip, int32_t n, uint32_t* out);
/* Boost >= 1.47 uniform_int<int>(lo, hi) draws on mt19937(seed). */
int orc_uniform_int(uint32_t seed, int32_t lo, int32_t hi, int32_t n, int32_t* out);
/* Same, but on an explicit stream of raw engine outputs (tempered words);
 * returns the number of words consumed in *used. */
int orc_uniform_int_words(const uint32_t* words, int32_t nwords, int32_t lo,
                          int32_t hi, int32_t n, int32_t* out, int32_t* used);

/* ---- stateless pair sums (Potential::coulomb* / gauss* / coulombImage*) ---- */
int orc_pair_potential(int32_t gx, int32_t gy, int32_t gz, int32_t cutoff,
                       double prefactor, double sigma,
                       const int32_t* charge_site, const int32_t* charge_q,
                       int32_t n_charges, const int32_t* target_site,
                       int32_t n_targets, int32_t flags, double* out);
/* FP64 closed form of the full-lattice map (kernel.cl coulomb1/gauss1):
 * out[s] = pref * sum_j q_j f(r) over 0 < r^2 < cutoff^2 (cutoff <= 0: none). */
int orc_potential_map(int32_t gx, int32_t gy, int32_t gz, int32_t cutoff,
                      double prefactor, double sigma,
                      const int32_t* charge_site, const int32_t* charge_q,
                      int32_t n_charges, int32_t flags, int32_t threads,
                      double* out);

/* ---- world ------------------------------------------------------------------ */
/* Builds a World exactly in the reference's construction order.  Explicit
 * lists play the role of checkpoint sections ([Electrons] [Holes] [Defects]
 * [Traps] [TrapPotentials]); pass n = 0 for none. */
orc_world* orc_world_create(const orc_params* p,
                            const int32_t* electrons, int32_t n_e,
                            const int32_t* holes, int32_t n_h,
                            const int32_t* defects, int32_t n_d,
                            const int32_t* traps, int32_t n_t,
                            const double* trap_potentials, int32_t n_tp);
void orc_world_destroy(orc_world* w);

/* performIterations(n) with the CPU Coulomb path; threads <= 1 is serial. */
int orc_world_step(orc_world* w, int32_t n, int32_t threads);
/* the three phases of one step, for phase-level parity checks */
int orc_world_phase_propose(orc_world* w);                 /* chooseFuture, all */
int orc_world_phase_coulomb(orc_world* w, int32_t threads); /* coulombCPU, all  */
int orc_world_phase_finish(orc_world* w);  /* decide, recombine, tick, inject   */

/* info[0..11] = n_e, n_h, n_d, max_e, max_h, V, X, Y, Z, special_e, special_h, n_traps */
int orc_world_info(orc_world* w, int32_t* info);
int64_t orc_world_current_step(orc_world* w);
/* which: 0 = electrons, 1 = holes; arrays sized by orc_world_info */
int orc_world_get_carriers(orc_world* w, int32_t which, int32_t* site,
                           int32_t* fsite, int32_t* lifetime,
                           int32_t* pathlength, double* de);
int orc_world_set_carriers(orc_world* w, int32_t which, int32_t n,
                           const int32_t* site, const int32_t* lifetime,
                           const int32_t* pathlength);
int orc_world_get_defects(orc_world* w, int32_t* site);
/* background potential of the lattice sites (size V), identical on both grids */
int orc_world_get_potential(orc_world* w, double* v);
/* agent types of both grids, size V + 35 each */
int orc_world_get_types(orc_world* w, uint8_t* etype, uint8_t* htype);
/* engine state as the window of the last 624 raw words and the number of
 * tempered outputs consumed since seeding */
int orc_world_get_rng(orc_world* w, uint32_t* window, uint64_t* consumed);
int orc_world_set_rng(orc_world* w, const uint32_t* window, uint64_t consumed);
/* 10 flux agents in World order: eSL eSR hSL hSR xS eDL eDR hDL hDR xD */
int orc_world_get_flux(orc_world* w, uint64_t* attempts, uint64_t* successes,
                       double* rates);
int orc_world_set_flux(orc_world* w, const uint64_t* attempts,
                       const uint64_t* successes);
/* derived constants: [prefactor, inverse_kt, sI[1][0][0], coupling(1,0,0)] */
int orc_world_constants(orc_world* w, double* out);
/* coupling table, (h+1)^3 doubles, index (dx*(h+1)+dy)*(h+1)+dz */
int orc_world_coupling(orc_world* w, double* out);
/* Per-carrier record of the last decideFuture call: the uniform draw u (or -1
 * when no draw happened) and the Metropolis margin
 *   m = coupling*exp(-beta*pd) - u   (pd > 0)   or   coupling - u  (pd <= 0)
 * with pd = q*(V(f)-V(s)) + dE; the move is accepted iff m > 0. */
int orc_world_last_decisions(orc_world* w, int32_t which, double* u,
                             double* margin, double* pd);

/* Grid::neighborsSite for one site, with one drain slot per x face (-1: none);
 * out needs 34 entries. */
int orc_neighbors(int32_t gx, int32_t gy, int32_t gz, int32_t h, int32_t site,
                  int32_t drain_l, int32_t drain_r, int32_t* out, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* LMORACLE_H */
