/*
 * CPU oracle for the adaptive SPH step -- TEST INFRASTRUCTURE ONLY.
 *
 * A phase-by-phase restatement, in plain C++, of the reference's CUDA kernels
 * (SPH-2D/, SPH-3D-2/, SPH-3D-4/ sph_system.cu and sph_system.cpp). It runs in two
 * precisions from one template: fp32 (the bit-exact parity checker and the CPU speed
 * baseline) and fp64 (the drift oracle). Only tests/, the smoke check and the CPU
 * baseline leg of the benchmark may load it; the product library never links it.
 *
 * Every phase cites the reference lines it restates in sph_oracle.cpp.
 *
 * Parity pins (tests/test_oracle_*.py): closed forms for hash/cell, pair and self density,
 * pressure, two-particle force, integrate (gravity, walls, spheres), uniform-flow energy,
 * isolated-particle NaN energy, split counts/offsets, merge centroids and counts, stable
 * compaction, single-particle density-grid splat, linear-field normals, sphere-SDF closed
 * marching-cubes mesh, brute-force O(N^2) density/force/energy on small random scenes, the
 * fp32 exp routine against libm double exp, and fp32 vs fp64 agreement.
 * The reference has no tests, fixtures or golden vectors of its own, and its CUDA build
 * cannot be compiled here, so bit-level agreement with the 2012 binary (CUDA pow/sqrt
 * rounding, cutil_math operator forms, thrust tie order) is "parity unpinned"; parity is
 * defined against this restatement.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* variants */
#define ORC_V2D 0  /* SPH-2D */
#define ORC_V3D2 1 /* SPH-3D-2 (1->2 split) */
#define ORC_V3D4 2 /* SPH-3D-4 (1->4 split) */

/* step flags */
#define ORC_SPLIT 1u
#define ORC_MERGE 2u
#define ORC_MC 4u

/* status codes (shared meaning with the product's C ABI) */
#define ORC_OK 0
#define ORC_EINVAL -1
#define ORC_EEMPTY -2
#define ORC_EDOMAIN -3
#define ORC_ECAP -4
#define ORC_ETRI -6

/* u32 accessors */
#define ORC_HASH 0       /* hash[k], sorted after orc_sort  (N)       */
#define ORC_INDEX 1      /* index[k], sorted after orc_sort (N)       */
#define ORC_START 2      /* cell_start (C)                            */
#define ORC_END 3        /* cell_end (C)                              */
#define ORC_SPLITFLAG 4  /* decide flag per slot, 0 = split (N)       */
#define ORC_STATUS 5     /* merge status per slot, 1 = dead (N')      */
/* f32 accessors */
#define ORC_DGRID 10     /* density grid (G)                          */
#define ORC_NORMAL 11    /* normals (3G)                              */
#define ORC_TRI 12       /* triangles: 18 floats each (p0 p1 p2 n0 n1 n2) */

typedef struct OrcSim OrcSim;

/* fp64: 0 = float32 (bit-exact checker), 1 = float64 (drift oracle), 2 and 3 = the
 * float64 and float32 builds with forward-mode derivatives (dual.h): every value is that
 * build's, and each carries its float64 derivative along one direction of the field and the
 * initial state */
OrcSim *orc_create(int variant, const void *sysparam, uint32_t capacity, int fp64);
void orc_destroy(OrcSim *s);
void orc_set_threads(int n);
int orc_get_threads(void);

int orc_set_params(OrcSim *s, const void *sysparam);
int orc_set_particles(OrcSim *s, const void *aos, uint32_t n);
int orc_get_particles(OrcSim *s, void *aos); /* writes num_particles records */
uint32_t orc_num_particles(OrcSim *s);
/* fp64 view of the state: 17 doubles per particle:
 * pos3 vel3 ev3 acc3 dens press energy surface level */
int orc_get_state_f64(OrcSim *s, double *out);

/* Differentiable-rollout model (no reference counterpart): a learnable force field of
 * nf^3 nodes x 3 components over the box, added to the force in integrate (nf = 0: off),
 * and the density of the particles at an ns^3 grid of sample points (ns^3 doubles). */
int orc_set_field(OrcSim *s, const float *theta, int nf);
int orc_set_field_f64(OrcSim *s, const double *theta, int nf); /* node values kept as doubles
                                                                  by the fp64 build */
int orc_density_sample(OrcSim *s, int ns, double *out);

/* Derivative builds only (fp64 = 2, 3): the field and its direction (nf^3 x 3 doubles each),
 * the direction of the initial state (3n doubles each for pos, vel, ev in slot order,
 * NULL = 0; call after orc_set_particles), and the density sample with its derivative. */
int orc_set_field_tangent(OrcSim *s, const double *theta, const double *dtheta, int nf);
int orc_set_state_tangent(OrcSim *s, const double *dpos, const double *dvel, const double *dev);
int orc_density_sample_tangent(OrcSim *s, int ns, double *value, double *tangent);

/* Rollout topology: mode 1 records each orc_step's split parents, merge groups and (3D)
 * branch decisions -- wall clamps, colour-field force on or off -- (the record is cleared),
 * mode 2 replays the record from its first step instead of deciding, mode 0 neither. orc_topology_copy copies the record of src into dst (any builds);
 * orc_topology_stats: {recorded steps, split parents, merge groups}. */
int orc_topology(OrcSim *s, int mode);
int orc_topology_copy(OrcSim *dst, const OrcSim *src);
int orc_topology_stats(const OrcSim *s, uint32_t out[3]);

int orc_calc_hash(OrcSim *s);
int orc_sort(OrcSim *s);
int orc_cell_ranges(OrcSim *s);
int orc_density(OrcSim *s);
int orc_force(OrcSim *s);
int orc_integrate(OrcSim *s);
int orc_energy(OrcSim *s);
int orc_split(OrcSim *s, uint32_t *n_requested, uint32_t *n_done);
int orc_merge(OrcSim *s, uint32_t *groups);
int orc_compact(OrcSim *s);
int orc_density_grid(OrcSim *s);
int orc_normals(OrcSim *s);
int orc_marching_cubes(OrcSim *s, uint32_t *n_tri);
int orc_step(OrcSim *s, uint32_t flags);

int orc_get_u32(OrcSim *s, int which, uint32_t *out, size_t n);
int orc_get_f32(OrcSim *s, int which, float *out, size_t n);
size_t orc_count(OrcSim *s, int which);

/* the fp32 exp/pow used by the wavelet (same published method as the product) */
float orc_expf(float t);
float orc_wavelet_f32(float x);
double orc_wavelet_f64(double x);

#ifdef __cplusplus
}
#endif
#endif
