/*
 * sph.h — C ABI of libsph, the MI355X (gfx950) SPH fluid solver.
 *
 * The solver replaces the OpenCL simulation path of paivett/fluids: the
 * abstract C++ class FluidSimulation (fluid/simulation/fluidsimulation.h:23-92)
 * with its two implementations PCISPHSimulation and WCSPHSimulation, the
 * uniform Grid, the BoundaryHandler and the OpenCL kernels under kernels/.
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (SPH_OK == 0, negative on error) unless stated otherwise, and
 * never lets a C++ exception cross the ABI; sph_last_error() gives the
 * message of the calling thread's last failure.
 *
 * Threading: a handle owns one HIP stream and its device buffers. A handle is
 * not thread-safe; different handles are independent (no global state besides
 * the per-thread error string).
 *
 * Particle order: the reference integrates into its "unsorted" buffers at the
 * sorted index (pcisphsimluation.cpp:160-162), so the stored order after a
 * step is the previous step's Morton order. sph_get_positions and friends
 * return that stored order, exactly as the reference's VBO holds it.
 */
#ifndef SPH_H_
#define SPH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_ERR_INVALID (-1)   /* bad argument */
#define SPH_ERR_HIP (-2)       /* HIP runtime failure (message holds the HIP error) */
#define SPH_ERR_STATE (-3)     /* call not valid in the current state */
#define SPH_ERR_RANGE (-4)     /* size or index out of range */
#define SPH_ERR_PARSE (-5)     /* configuration file could not be parsed */

/* SimulationSettings::Method (settings/simulationsettings.h:28-31) */
#define SPH_WCSPH 0
#define SPH_PCISPH 1

/* sph_settings.flags. 0 reproduces the reference exactly. */
#define SPH_FIX_WCSPH_DENSITY 1  /* WCSPH density evaluates W_POLY6 on r^2 instead of r (wcsph.cl:61-62,70-71) */
/*
 * Single-scene solvers: the fluid kernels divide and take square roots with
 * the GPU's approximate reciprocal, square-root and reciprocal-square-root
 * instructions (about 1 ulp each), as the reference builds its kernels with
 * -cl-fast-relaxed-math (pcisphsimluation.cpp:496-497,
 * wcsphsimluation.cpp:364). Neighbour search, cell keys and the boundary
 * wrench are unchanged; fields differ from the default IEEE build by rounding
 * only. Batches (sph_batch_add_scene) and the adjoint reject it.
 */
#define SPH_FAST_MATH 2
/*
 * Single-scene solvers: sort every step's cell keys with the full radix sort.
 * By default a step sorts only the keys that changed since the previous
 * step's sorted order and merges them in (same stable order, bit for bit);
 * this flag is for comparing the two.
 */
#define SPH_FULL_SORT 4

/*
 * PhysicsSettings (settings/physicssettings.h:5-25) and SimulationSettings
 * (settings/simulationsettings.h:9-39) merged into one plain struct; the keys
 * are those of config.pci / config.wc.
 */
typedef struct sph_settings {
    float time_step;
    float max_vel;
    float fluid_particle_radius;
    float fluid_support_radius;
    int pcisph_max_iterations;
    float pcisph_error_ratio;
    int method; /* SPH_WCSPH or SPH_PCISPH */
    float rest_density;
    float k_viscosity;
    float gravity;
    float gas_stiffness;
    float surface_tension;
    int flags; /* SPH_FIX_* bits; 0 = reference behaviour */
} sph_settings;

/* Per-step diagnostics (no counterpart in the reference, whose failure modes
 * are silent: list truncation at 50, grid wrap, out-of-range cell reads). */
typedef struct sph_step_stats {
    int iterations;               /* PCISPH corrections run in the last step */
    int fluid_list_overflows;     /* particles whose fluid list hit the cap of 50 */
    int boundary_list_overflows;  /* particles whose boundary list hit the cap of 50 */
    int edge_reads;               /* neighbour-cell ids outside [0, cps^3) (undefined reads in the reference) */
    int out_of_grid;              /* particles whose cell lies outside the allocated range box */
    int nan_positions;            /* particles with a non-finite position after the step */
    long long fluid_neighbours;   /* sum of fluid list lengths */
    long long boundary_neighbours;/* sum of boundary list lengths */
    float max_density_error;      /* max(rho* - rho0, 0) of the last correction */
} sph_step_stats;

typedef struct sph_handle sph_handle;

/* Message of the calling thread's last failure ("" if none). */
const char* sph_last_error(void);

/* Library version string. */
const char* sph_version(void);

/* ---------------------------------------------------------------- settings */

/* Reference defaults of PhysicsSettings/SimulationSettings
 * (physicssettings.h:17-22, simulationsettings.h:33-39). */
int sph_default_settings(sph_settings* out);

/* Settings::load (settings/settings.cpp:14-59, settings/fileparser.cpp:7-50):
 * key=value file, first whitespace token of each side, every simulation and
 * physics key required, unknown keys ignored, method must be wcsph|pcisph.
 * render_method (a GraphicsSettings key) is checked when present but not
 * required. Returns SPH_ERR_PARSE with the reference's message on a missing
 * file or key. */
int sph_load_settings(const char* path, sph_settings* out);

/* ------------------------------------------------------------ construction */

/* FluidSimulationFactory::build_simulation (fluidsimulationfactory.h:13-40):
 * constructs the solver selected by settings->method. Returns NULL on error. */
sph_handle* sph_create(const sph_settings* settings);

/* FluidSimulation::~FluidSimulation (fluidsimulation.h:29). */
int sph_destroy(sph_handle* h);

/* FluidSimulation::reset (fluidsimulation.h:47-48; pcisphsimluation.cpp:257-266,
 * wcsphsimluation.cpp:205-214): new parameters, boundary psi recomputed,
 * fluid re-initialised from the stored volumes (PCISPH runs its relaxation
 * step). */
int sph_reset(sph_handle* h, const sph_settings* settings);

/* FluidSimulation::set_rect_limits (fluidsimulation.h:79;
 * pcisphsimluation.cpp:557-564): the clamp box of the integrators. */
int sph_set_rect_limits(sph_handle* h, float width, float height, float depth);

/* FluidSimulation::add_volume (fluidsimulation.h:91; pcisphsimluation.cpp:525-530)
 * with a FluidVolume whose particles() (fluidvolume.h:10) returns this fixed
 * set. Re-initialises the fluid from all volumes; PCISPH then runs one
 * relaxation step and zeroes the velocities (pcisphsimluation.cpp:548-550). */
int sph_add_volume(sph_handle* h, const float* xyzw, int n);

/* FluidSimulation::add_volume with a BoxVolume (fluid/simulation/boxvolume.cpp:11-40),
 * re-sampled from the particle radius on every (re)initialisation. */
int sph_add_box_volume(sph_handle* h, const float size[3], const float center[3]);

/* FluidSimulation::add_boundary (fluidsimulation.h:66-68;
 * boundaryhandler.cpp:95-153) for a RigidBody whose surface_sampling()
 * (rigidbody.h:104) returned local_xyzw. R (row-major basis rows, as
 * btTransform::getBasis) and com give the body's transform at sampling time
 * (boundaryhandler.cpp:259-274). Only bodies with mass > 0 receive a fluid
 * wrench (boundaryhandler.cpp:322). Returns the body id (>= 0) or an error. */
int sph_add_boundary(sph_handle* h, const float* local_xyzw, int n, int can_move, float mass,
                     const float R[9], const float com[3]);

/* Replaces the RigidBody reads of BoundaryHandler::sync (boundaryhandler.cpp:259-270:
 * transform(), angular_vel(), linear_vel(), position()) for the next step. Any
 * pointer may be NULL to keep the previous value. */
int sph_set_body_state(sph_handle* h, int body, const float R[9], const float com[3],
                       const float lin_vel[3], const float ang_vel[3]);

/* Replaces RigidBody::apply_force_n_torque (rigidbody.cpp:49-54): the summed
 * fluid force and torque (about com) of the last step (clsum of
 * boundaryhandler.cpp:338-346). Zero for bodies with mass <= 0. */
int sph_get_body_wrench(sph_handle* h, int body, float force[3], float torque[3]);

/* ------------------------------------------------------------------- step */

/* FluidSimulation::simulate (fluidsimulation.h:36): one step with the
 * settings' time step (PCISPHSimulation::_simulate_pcisph_step,
 * pcisphsimluation.cpp:72-165; WCSPHSimulation::simulate,
 * wcsphsimluation.cpp:40-106). Asynchronous on the handle's stream. */
int sph_step(sph_handle* h);

/* sph_step repeated n times (no host synchronisation in between). */
int sph_step_n(sph_handle* h, int n);

/* step(dt): sets the time step (recomputing the PCISPH scale factor, which
 * depends on it: pcisphsimluation.cpp:432) and runs one step. */
int sph_step_dt(sph_handle* h, float dt);

/* Waits for the handle's stream. */
int sph_synchronize(sph_handle* h);

/* FluidSimulation::particle_count / boundary_particle_count
 * (fluidsimulation.h:54,84). Return the count (>= 0) or an error. */
int sph_particle_count(sph_handle* h);
int sph_boundary_particle_count(sph_handle* h);

/* --------------------------------------------------------- buffers (host) */

/* Positions/velocities in the stored order (xyzw, n = particle_count). These
 * replace the reference's GL VBO (scene/fluid.cpp:27) and its device-only
 * velocity buffers. sph_set_* write w = 1 for positions and w = 0 for
 * velocities, which is what every producer of the reference writes. */
int sph_get_positions(sph_handle* h, float* xyzw, int n);
int sph_set_positions(sph_handle* h, const float* xyzw, int n);
int sph_get_velocities(sph_handle* h, float* xyzw, int n);
int sph_set_velocities(sph_handle* h, const float* xyzw, int n);
/* WCSPH half-step velocities (wcsphsimluation.cpp:140). */
int sph_get_half_velocities(sph_handle* h, float* xyzw, int n);
int sph_set_half_velocities(sph_handle* h, const float* xyzw, int n);

/* Field ids for sph_get_field. Fields are left by the last step, in sorted
 * order unless stated; vector fields are 4 floats per element. */
#define SPH_FIELD_KEYS 0            /* u32[N]  sorted Morton cell keys (grid.cl:35) */
#define SPH_FIELD_PERM 1            /* i32[N]  sort permutation (the reference's mask) */
#define SPH_FIELD_NCOUNT 2          /* i32[N]  fluid neighbour counts */
#define SPH_FIELD_NLIST 3           /* i32[50N] fluid lists, column-major list[k*N+i], unused slots -1 */
#define SPH_FIELD_BNCOUNT 4         /* i32[N]  boundary neighbour counts */
#define SPH_FIELD_BNLIST 5          /* i32[50N] boundary lists */
#define SPH_FIELD_POS_SORTED 6      /* f4[N] */
#define SPH_FIELD_VEL_SORTED 7      /* f4[N] */
#define SPH_FIELD_DENSITY 8         /* f32[N]  initial density (PCISPH) / density (WCSPH) */
#define SPH_FIELD_PRESSURE 9        /* f32[N]  accumulated pressure (PCISPH) / Tait pressure (WCSPH) */
#define SPH_FIELD_ERR 10            /* f32[N]  density variation of the last correction */
#define SPH_FIELD_NORMAL 11         /* f4[N] */
#define SPH_FIELD_FORCE_OTHER 12    /* f4[N]  viscosity + surface tension force */
#define SPH_FIELD_FORCE_PRESSURE 13 /* f4[N]  pressure force of the last correction */
#define SPH_FIELD_POS_PRED 14       /* f4[N]  predicted positions of the last correction */
#define SPH_FIELD_ACCEL 15          /* f4[N]  WCSPH acceleration */
#define SPH_FIELD_POS 16            /* f4[N]  stored-order positions */
#define SPH_FIELD_VEL 17            /* f4[N]  stored-order velocities */
#define SPH_FIELD_VEL_HALF 18       /* f4[N]  stored-order WCSPH half-step velocities */
#define SPH_FIELD_BPOS_SORTED 19    /* f4[B]  sorted boundary positions */
#define SPH_FIELD_BPHI_SORTED 20    /* f32[B] sorted boundary psi */
#define SPH_FIELD_BKEYS 21          /* u32[B] sorted boundary keys */
#define SPH_FIELD_BPERM 22          /* i32[B] boundary sort permutation */
#define SPH_FIELD_BPHI 23           /* f32[B] psi in sampling order */
#define SPH_FIELD_BFORCE 24         /* f4[B]  per-boundary-particle fluid force (bodies with mass > 0) */
#define SPH_FIELD_BTORQUE 25        /* f4[B]  per-boundary-particle torque */
#define SPH_FIELD_RANGES 26         /* i32[3*C] (reference cell id, start, end) of occupied fluid cells, by start */
#define SPH_FIELD_BRANGES 27        /* i32[3*C] same for the boundary grid */
#define SPH_FIELD_VEL_HALF_SORTED 28/* f4[N] */
#define SPH_FIELD_COUNT 29

/* Copies field `field` into out (capacity `cap` elements of the field's
 * element type; pass out == NULL to query the element count). Synchronises
 * the stream. Returns the element count or an error. */
int sph_get_field(sph_handle* h, int field, void* out, long long cap);

/* Diagnostics of the last step (synchronises; one reduction pass over the
 * particles, about 0.05 ms at 1M, so it can run after every step). */
int sph_get_step_stats(sph_handle* h, sph_step_stats* out);

/* Device pointers of the stored-order buffers (float4[N]) for zero-copy use,
 * valid until the next call that changes the particle count. */
int sph_device_state(sph_handle* h, void** pos, void** vel, void** vel_half);
/* The handle's HIP stream (hipStream_t). */
void* sph_stream(sph_handle* h);

/* ------------------------------------------------------------- profiling */

/* Per-kernel HIP-event timing on the handle's stream (off by default). */
int sph_set_profiling(sph_handle* h, int on);
/* Restricts the timing to launches of one kernel (NULL or "" times all). */
int sph_set_profiling_filter(sph_handle* h, const char* kernel);
/* Fills names (a buffer of `cap` * 48 chars), launches and total milliseconds
 * per kernel name; returns the number of kernel names. */
int sph_get_kernel_times(sph_handle* h, char* names, int* launches, double* total_ms, int cap);
int sph_reset_kernel_times(sph_handle* h);

/* ------------------------------------------------------------ autotuning */

/* Launch shapes of a PCISPH handle's list kernels, the single-scene form of
 * the reference's launch autotuning (CLKernel::run, opencl/clkernel.cpp:
 * 85-151, which times work-group sizes during the first launches). Kernels:
 * SPH_TUNE_* below. A shape is threads per block (64, 128 or 256) and, for
 * update_pressure and pressure_force, a fixed grid of blocks (0 = one thread
 * per particle). Shapes change timing only, never results. SPH_ERR_STATE for
 * a WCSPH handle. */
#define SPH_TUNE_NORMALS 0
#define SPH_TUNE_FORCES_PREDICT 1
#define SPH_TUNE_UPDATE_PRESSURE 2
#define SPH_TUNE_PRESSURE_FORCE 3
/* Times every candidate shape of each kernel over `trials` steps from the
 * current state, keeps the fastest and restores the state (positions and
 * velocities). cache_path (may be NULL) names a text file of earlier results
 * for the same device and scene size: when it holds them they are used, else
 * the new ones are appended. Returns 1 if the cache supplied the shapes, 0 if
 * they were measured. */
int sph_autotune(sph_handle* h, const char* cache_path, int trials);
int sph_get_launch(sph_handle* h, int kernel, int* block, int* grid);
int sph_set_launch(sph_handle* h, int kernel, int block, int grid);

/* ------------------------------------------------------------------ sort */

/* clsort<cl_uint, cl_int> (opencl/algorithms/clsort.h:12-35): stable sort of
 * n (key, value) pairs in place by the low key_bits bits of the keys (0 = all
 * 32, as clogs::Radixsort with maxBits 0), on device arrays the caller owns.
 * keys (uint32_t) and values (int32_t) are device pointers of n elements;
 * stream is a hipStream_t (NULL = the null stream). Asynchronous on that
 * stream: the library takes its work arrays (8 bytes per pair plus the digit
 * histograms) from the stream-ordered allocator and returns them on the same
 * stream. n up to 2^27. */
int sph_sort_pairs(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream);

/* No counterpart in the reference: sph_sort_pairs for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, same result bit for bit; its work arrays
 * live in a fixed array of the library's device image (about 10 MB, held by
 * every GPU the library runs on until the process ends), so it calls no
 * allocator. n is limited to SPH_SORT_NOALLOC_MAX (SPH_ERR_RANGE beyond).
 * Every call on one GPU shares that array: calls that overlap in time on one
 * GPU (other streams, threads or handles) race and must be serialised by the
 * caller. */
#define SPH_SORT_NOALLOC_MAX (1 << 20)
int sph_sort_pairs_noalloc(uint32_t* keys, int32_t* values, int n, int key_bits, void* stream);

/* The stable sort of (keys[i], i) for keys that differ little from a sorted
 * array: the sort the single-scene solvers run every step, where particles
 * are stored in the previous step's sorted order (replacing the per-step
 * clsort of pcisphsimluation.cpp:86 / wcsphsimluation.cpp:54 by one
 * that places only the keys that changed). hint: n non-decreasing keys (any
 * such array gives the exact result; the closer to sorted(keys), the less
 * work). keys: the n keys to sort. sorted_keys receives them in sorted order
 * and perm the stable permutation (perm[r] = index of the r-th key), the same
 * bits as sph_sort_pairs on (keys, 0 .. n-1). All device arrays; sorted_keys
 * may be hint itself, not keys. Work arrays come from the stream-ordered
 * allocator. Checks the hint (SPH_ERR_INVALID if it decreases) and therefore
 * synchronises the stream. n up to 2^27. */
int sph_sort_near(const uint32_t* hint, const uint32_t* keys, uint32_t* sorted_keys, int32_t* perm, int n,
                  void* stream);

/* -------------------------------------------------------------- samplers */
/* Each returns the sample count; out (xyzw) receives up to cap samples. */

/* BoxVolume::particles (fluid/simulation/boxvolume.cpp:11-40). */
int sph_sample_box_volume(const float size[3], const float center[3], float radius, float* out, int cap);
/* FishTank::surface_sampling (scene/fishtank.cpp:108-141), tank at the origin. */
int sph_sample_fishtank(float width, float height, float depth, float spacing, float* out, int cap);
/* Cube::surface_sampling (scene/cube.cpp:136-170), body-local. */
int sph_sample_cube(float side, float spacing, float* out, int cap);
/* Wall::surface_sampling (scene/wall.cpp:142-152), body-local. */
int sph_sample_wall(float width, float height, float spacing, float* out, int cap);
/* Sphere::surface_sampling (scene/sphere.cpp:135-174) over IcoSphereMesh(radius, 3), body-local. */
int sph_sample_sphere(float radius, float spacing, float* out, int cap);
/* Model::surface_sampling (scene/model.cpp:170-214): voxel corners of an OBJ
 * mesh (mesh/modelmesh.cpp) at the given spacing, body-local. */
int sph_sample_model(const char* obj_path, float spacing, float* out, int cap);

/* ------------------------------------------------------- rigid bodies (F1)
 * A host rigid-body world in place of Bullet's btDiscreteDynamicsWorld for
 * the two-way coupling loop (scene/scene.cpp:100-110, 304-316): gravity
 * (0, -9.8, 0) as the reference hard-wires it, one variable step of dt per
 * frame, contacts with the container's floor and walls and between bodies.
 * Quaternions are (x, y, z, w) like btQuaternion and need not be unit. */
typedef struct sph_rigid_world sph_rigid_world;

/* btDiscreteDynamicsWorld + setGravity(0, gravity_y, 0) (scene.cpp:304-316). */
sph_rigid_world* sph_rigid_create(float gravity_y);
int sph_rigid_destroy(sph_rigid_world* w);
/* FishTank's collision mesh (scene/fishtank.cpp:143-182): an open-top box of
 * the given sides centred at center (NULL = origin). */
int sph_rigid_set_container(sph_rigid_world* w, float width, float height, float depth, const float center[3]);
/* Cube::_init_physics (scene/cube.cpp:129-134): a box body of the given
 * half extents; mass 0 makes it static. Returns the body id. */
int sph_rigid_add_box(sph_rigid_world* w, float mass, const float half[3], const float pos[3], const float quat[4]);
/* Sphere's btSphereShape body (scene/sphere.cpp). Returns the body id. */
int sph_rigid_add_sphere(sph_rigid_world* w, float mass, float radius, const float pos[3]);
/* A body that never moves and collides with nothing (Wall, Model). */
int sph_rigid_add_static(sph_rigid_world* w, const float pos[3], const float quat[4]);
/* RigidBody::apply_force_n_torque (scene/rigidbody.cpp:49-54): clearForces,
 * applyCentralForce(f), applyTorque(t). */
int sph_rigid_apply_force_n_torque(sph_rigid_world* w, int body, const float f[3], const float t[3]);
/* btDiscreteDynamicsWorld::stepSimulation(dt, 0) (scene.cpp:109). */
int sph_rigid_step(sph_rigid_world* w, float dt);
/* RigidBody::transform / position / rotation / linear_vel / angular_vel
 * (scene/rigidbody.cpp:27-62): basis rows R, centre of mass, quaternion,
 * velocities. Any output may be NULL. */
int sph_rigid_get_state(sph_rigid_world* w, int body, float R[9], float com[3], float quat[4], float lin_vel[3],
                        float ang_vel[3]);
/* RigidBody::set_position / set_rotation, and velocities (NULL = keep). */
int sph_rigid_set_state(sph_rigid_world* w, int body, const float pos[3], const float quat[4], const float lin_vel[3],
                        const float ang_vel[3]);
int sph_rigid_body_count(sph_rigid_world* w);

/* One frame of Scene::render (scene/scene.cpp:100-110) for bodies that are
 * both fluid boundaries and rigid bodies: for each pair k, the body state of
 * rigid body rigid_ids[k] goes to fluid body sph_ids[k]
 * (BoundaryHandler::sync, boundaryhandler.cpp:259-270); the fluid steps; each
 * pair whose rigid body has mass > 0 gets the fluid's force and torque
 * (apply_fluid_forces, boundaryhandler.cpp:322-346); then the rigid world
 * steps by the fluid's time step. Synchronises the fluid's stream. */
int sph_coupled_step(sph_handle* h, sph_rigid_world* w, int npairs, const int* sph_ids, const int* rigid_ids);

/* ------------------------------------------------------ batched WCSPH (C4)
 * Many independent WCSPH scenes advanced together: the scenes' particles are
 * packed into one array and every step is one set of launches over all of
 * them (captured once into a HIP graph and replayed), with per-scene grids
 * and parameters. A scene in a batch produces the same bits as the same
 * scene run alone through sph_create / sph_step. No reference counterpart:
 * the reference runs one scene per process (fluidsimulationfactory.h:13-40);
 * callers are sweeps and ensembles of small scenes. Scenes have no boundary
 * particles (their tanks are the integrator's clamp box, Q15). */
typedef struct sph_batch sph_batch;

sph_batch* sph_batch_create(void);
int sph_batch_destroy(sph_batch* b);
/* A scene: WCSPH settings, its container (as sph_set_rect_limits) and its
 * particles (xyzw, as sph_add_volume). Returns the scene id. */
int sph_batch_add_scene(sph_batch* b, const sph_settings* s, float width, float height, float depth,
                        const float* xyzw, int n);
/* Packs the scenes and allocates device memory; no scene can be added after. */
int sph_batch_finalize(sph_batch* b);
/* Times the launch shapes of each batched kernel (threads per block, blocks,
 * LDS-staged list slots) on this batch and keeps the fastest; the scenes'
 * state is restored afterwards. cache_path (may be NULL) names a text file of
 * earlier results for the same device and batch size. Returns 1 if the cache
 * supplied the shapes, 0 if they were measured. */
int sph_batch_autotune(sph_batch* b, const char* cache_path, int trials);
/* kernel: 0 hash, 1 gather+ranges, 2 lists+density, 3 normals, 4 accel+integrate */
int sph_batch_get_launch(sph_batch* b, int kernel, int* block, int* grid, int* stage);
int sph_batch_set_launch(sph_batch* b, int kernel, int block, int grid, int stage);
/* Graph replay of the step (on by default). */
int sph_batch_set_graph(sph_batch* b, int on);
int sph_batch_step(sph_batch* b, int nsteps);
int sph_batch_synchronize(sph_batch* b);
int sph_batch_scene_count(sph_batch* b);
int sph_batch_particle_count(sph_batch* b, int scene);
int sph_batch_total_particles(sph_batch* b);
/* Stored-order float4 state of one scene (synchronises). */
int sph_batch_get_positions(sph_batch* b, int scene, float* xyzw, int n);
int sph_batch_get_velocities(sph_batch* b, int scene, float* xyzw, int n);
int sph_batch_get_half_velocities(sph_batch* b, int scene, float* xyzw, int n);
int sph_batch_set_positions(sph_batch* b, int scene, const float* xyzw, int n);
int sph_batch_set_velocities(sph_batch* b, int scene, const float* xyzw, int n);
/* Last step's densities (float) and neighbour counts (int) of one scene, in
 * its sorted order. */
int sph_batch_get_density(sph_batch* b, int scene, float* out, int n);
int sph_batch_get_neighbour_counts(sph_batch* b, int scene, int* out, int n);
/* HIP-event timing of the batched launches, as sph_set_profiling. */
int sph_batch_set_profiling(sph_batch* b, int on, const char* kernel);
int sph_batch_get_kernel_times(sph_batch* b, char* names, int* launches, double* total_ms, int cap);

/* --------------------------------------------------- WCSPH adjoint (C5)
 * Gradients of a loss on the state after T WCSPH steps with respect to the
 * initial state and to k_viscosity, gas_stiffness and surface_tension,
 * computed by a hand-written reverse-mode pass. The forward pass is the
 * solver's own step (same kernels, same bits as sph_step) with a checkpoint
 * every C steps; the backward pass recomputes each segment from its
 * checkpoint and runs the adjoint kernels backwards through it. No reference
 * counterpart (the reference has no gradients); callers are inverse problems
 * and parameter fits, e.g. through fluids_amd.adjoint.WCSPHRollout, a
 * torch.autograd.Function. Scenes have no boundary particles and every
 * neighbour list must stay below the cap of 50 (else the backward pass
 * returns SPH_ERR_RANGE). States are float4 per particle in particle-id
 * order. */
typedef struct sph_adjoint sph_adjoint;

sph_adjoint* sph_adjoint_create(const sph_settings* s, float width, float height, float depth, const float* pos,
                                const float* vel, const float* vel_half, int n, int checkpoint_every);
int sph_adjoint_destroy(sph_adjoint* a);
int sph_adjoint_set_initial(sph_adjoint* a, const float* pos, const float* vel, const float* vel_half);
int sph_adjoint_set_params(sph_adjoint* a, float k_viscosity, float gas_stiffness, float surface_tension);
int sph_adjoint_forward(sph_adjoint* a, int nsteps);
/* State after the forward pass (any output may be NULL). */
int sph_adjoint_final_state(sph_adjoint* a, float* pos, float* vel, float* vel_half);
/* g_* : dL/d(final pos, vel, vel_half) (NULL = 0). Outputs (may be NULL):
 * dL/d(initial pos, vel, vel_half) and g_params = dL/d(k_viscosity,
 * gas_stiffness, surface_tension). */
int sph_adjoint_backward(sph_adjoint* a, const float* g_pos, const float* g_vel, const float* g_vel_half,
                         float* g_pos0, float* g_vel0, float* g_vel_half0, double* g_params);
/* Wall-clock milliseconds of the last forward and backward passes. */
int sph_adjoint_timings(sph_adjoint* a, double* forward_ms, double* backward_ms);

/* ------------------------------------------------------ scene files (F2)
 * Scene::load_scene (scene/scene.cpp:322-476) without rendering: reads a
 * scene file (container, box fluid volumes, cube / sphere / model / wall
 * bodies), builds the solver in the reference's init order, samples each
 * body at twice the particle radius (boundaryhandler.cpp:19) as a boundary
 * that moves when its mass is > 0, and puts the bodies and a FishTank
 * container into a rigid world. base_dir resolves relative model paths. */
typedef struct sph_scene sph_scene;

sph_scene* sph_scene_load(const char* path, const sph_settings* s, const char* base_dir);
int sph_scene_destroy(sph_scene* sc);
/* The scene's solver and rigid world, owned by the scene (do not destroy). */
sph_handle* sph_scene_simulation(sph_scene* sc);
sph_rigid_world* sph_scene_world(sph_scene* sc);
/* Frames of Scene::render (scene.cpp:100-110): sph_coupled_step over every
 * body with mass > 0. */
int sph_scene_step(sph_scene* sc, int nsteps);
int sph_scene_body_count(sph_scene* sc);
/* Body k: type (16 chars), name (64 chars), its fluid boundary id, its rigid
 * body id, mass and sample count. Outputs may be NULL. */
int sph_scene_body(sph_scene* sc, int k, char* type, char* name, int* sph_id, int* rigid_id, float* mass,
                   int* samples);

/* ------------------------------------------------------------ output (F4)
 * In place of the GL vertex buffer the reference renders from
 * (opencl/clallocator.h:165-183, 300-374) the particle buffers leave the
 * device as a frame file, and in place of the -o option (main.cpp:19,
 * 34-35; gui/glwidget.cpp:48-73; gui/mainwindow.cpp:99-101) frame times go
 * to an FPS log.
 *
 * Frame file: "SPHF", u32 version (1), u32 fields (bit 0 positions, bit 1
 * velocities), u32 0; per frame: u64 frame index, f64 simulated time, u32 n,
 * u32 fields, n float4 positions then n float4 velocities (as present), in
 * stored order. Reader: fluids_amd.output.read_frames. */
typedef struct sph_frames sph_frames;
typedef struct sph_perf_log sph_perf_log;

sph_frames* sph_frames_open(sph_handle* h, const char* path, int fields);
/* Queues the current state: an asynchronous device-to-pinned copy on the
 * handle's stream; a background thread writes it. */
int sph_frames_write(sph_frames* f);
/* Writes the queued frames, closes the file, frees f; returns the frame count. */
int sph_frames_close(sph_frames* f);
/* FPS log: call sph_perf_frame after each frame (it waits for the handle's
 * stream); every 6 frames one line with the frames-per-second of those
 * frames is written, as the reference's widget does. */
sph_perf_log* sph_perf_open(sph_handle* h, const char* path);
int sph_perf_frame(sph_perf_log* p);
/* Closes the log, frees p; returns the number of lines written. */
int sph_perf_close(sph_perf_log* p);

#ifdef __cplusplus
}
#endif

#endif /* SPH_H_ */
