/*
 * mdacp_oracle.h -- CPU restatement of MDACP's single-domain Lennard-Jones step.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path's results are tested against this library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (mdacp_amd/libmdgpu.so) never links, loads or calls it.
 *
 * Every function restates one reference routine (file:line in kaityo256/mdacp)
 * with the same floating-point operation order, so that results are bit-exact
 * with the reference's default (-O3, no FMA) build on the same inputs.
 */
#ifndef MDACP_ORACLE_H
#define MDACP_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_sys ora_sys;

/* error codes returned by the stepping functions (same meaning as mdgpu.h) */
#define ORA_OK 0
#define ORA_ERR_TOO_FAST 1        /* pairlist.cc:58-62 */
#define ORA_ERR_INVALID_POSITION 2 /* variables.cc:70-78 */
#define ORA_ERR_INVALID_INDEX 3    /* meshlist.cc:178-182 */

/* thermostat selectors (mdconfig.h:27, simulationinfo.cc:17-25) */
#define ORA_NVE 0
#define ORA_NOSEHOOVER 1
#define ORA_LANGEVIN 2
/* Langevin noise source: the reference's mt19937 + normal_distribution stream,
   or the counter-based Philox4x32-10 stream that the GPU path uses */
#define ORA_NOISE_MT19937 0
#define ORA_NOISE_PHILOX 1

ora_sys *ora_create(double Lx, double Ly, double Lz, double dt, double rc, double margin);
void ora_destroy(ora_sys *s);
void ora_set_thermostat(ora_sys *s, int kind, double T, double gamma, double tau);
void ora_set_noise(ora_sys *s, int noise_kind, uint64_t philox_seed);
void ora_set_check_list_length(ora_sys *s, int n);
int ora_add_particles(ora_sys *s, int n, const double *x, const double *v, const int *type);
void ora_set_initial_velocity(ora_sys *s, double v0, int seed);
void ora_set_seed(ora_sys *s, int seed);

int ora_make_pair_list(ora_sys *s);      /* MDManager::MakePairList */
int ora_calculate(ora_sys *s);           /* MDManager::Calculate */
int ora_last_step_rebuilt(const ora_sys *s);
void ora_send_border_particles(ora_sys *s);
void ora_change_scale(ora_sys *s, double alpha); /* MDManager::ChangeScale */
void ora_update_position_half(ora_sys *s);
void ora_calculate_force(ora_sys *s);     /* kick only (CalculateForceNext) */
void ora_kick_only(ora_sys *s, double *dp_out); /* dp = kick from p=0, owned rows */

double ora_kinetic_energy_raw(ora_sys *s);   /* KineticEnergyObserver */
double ora_potential_energy_raw(ora_sys *s); /* PotentialEnergyObserver */
double ora_virial_raw(ora_sys *s);           /* VirialObserver */
double ora_temperature(ora_sys *s);
double ora_pressure(ora_sys *s);
double ora_total_energy(ora_sys *s);
double ora_sim_time(const ora_sys *s);
double ora_zeta(const ora_sys *s);

int ora_particle_number(const ora_sys *s);
int ora_total_particle_number(const ora_sys *s);
long ora_pair_number(const ora_sys *s);
int ora_mesh_dims(const ora_sys *s, int *m3);
void ora_get_state(const ora_sys *s, double *q, double *p, int *type);
void ora_set_momenta(ora_sys *s, const double *p); /* by owned index */
void ora_get_ghosts(const ora_sys *s, int *src_type, int *shift_code, double *q);
long ora_get_pairs(const ora_sys *s, int *out3); /* (id_a, id_b, shift) per COO pair */
void ora_get_cells(const ora_sys *s, int *cell_of_owned);
double ora_buffer_length(const ora_sys *s);
int ora_lifetime(const ora_sys *s);

/* standalone pieces of the reference's third-party arithmetic */
void ora_mt_raw(uint32_t seed, int n, uint32_t *out);                /* std::mt19937 outputs */
void ora_mt_uniform(uint32_t seed, int n, double *out);            /* mt19937 + uniform_real(0,1) */
void ora_mt_normal(uint32_t seed, double sd, int n, double *out);   /* normal_distribution(0,sd) */
void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ora_philox_normal3(uint64_t seed, uint32_t id, uint64_t step, double out[3]);

/* FCC lattice of confmaker.cc:8-44 for an arbitrary box; returns count */
long ora_fcc(double Lx, double Ly, double Lz, double density, long cap, double *x);

#ifdef __cplusplus
}
#endif
#endif
