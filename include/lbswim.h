/*
 * lbswim C ABI -- the drop-in boundary of the MI355X-native coupled
 * lattice-Boltzmann / swimmer / tracer step.
 *
 * The reference exposes its library to Python through Cython `cdef extern`
 * declarations of C++ classes (lbswim/_lb.pxd:3-49, lbswim/_shared.pxd:1-13,
 * lbswim/SwimmerArray.pxd:4-26, lbswim/TracerArray.pxd:4-11).  Each entry
 * point below replaces one member of that surface; the comment on each names
 * it.  Plain C types only: opaque handles, pointers, sizes and int status codes.
 *
 * Conventions
 *   - Every function returning int returns LBSWIM_OK (0) on success or a
 *     positive LBSWIM_E* code; lbswim_last_error() then describes the failure
 *     (per thread).  The reference prints and calls exit() instead
 *     (lbswim/liblbswim/cucall.h:5-14).
 *   - Data layout is the reference's: lattice site ijk = i*ny*nz + j*nz + k;
 *     fields element-slowest, f[p*n + ijk], u[a*n + ijk]; particle vectors
 *     r[a*N + i] (README.md:39; lb.cu:49-57,117).
 *   - Every object owns a host mirror and a device copy of each field
 *     (SharedArray, lbswim/liblbswim/Shared.h:4-25).  Host mirrors are FP64
 *     (the particle PRNG mirror is uint64) and zero-initialised; *_h2d and
 *     *_d2h copy a whole field and return when the copy is complete.
 *   - Device work of all objects is ordered on one library stream, as the
 *     reference's default stream orders it; lbswim_synchronize() waits for it.
 *     Lattice/particle calls only enqueue (the reference's Step also waits,
 *     lb.cu:239).  Calls are not thread-safe.
 */
#ifndef LBSWIM_H
#define LBSWIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBSWIM_Q 15 /* DQ_q, lbswim/liblbswim/d3q15.h:5 */
#define LBSWIM_D 3  /* DQ_d, lbswim/liblbswim/d3q15.h:4 */

/* status codes */
#define LBSWIM_OK 0
#define LBSWIM_EINVAL 1 /* invalid argument (size, tau, count, field id, NULL) */
#define LBSWIM_ENOMEM 2 /* host or device allocation failed */
#define LBSWIM_EHIP 3   /* HIP runtime error (launch, copy, stream) */
#define LBSWIM_ENODEV 4 /* no usable GPU */

/* storage forms of the lattice fields */
#define LBSWIM_STORE_F64 0    /* FP64 everywhere: the reference's arithmetic */
#define LBSWIM_STORE_F32DEV 1 /* FP32 f as deviation from the rest state, FP32 rho/u/force */

/* how the coupled step hands the swimmer force to the lattice kernels */
#define LBSWIM_FORCE_STORED 0 /* written to the force array, read back by Step and CalcHydro */
#define LBSWIM_FORCE_AUTO 1   /* virtual where the swimmers are sparse (the default) */
#define LBSWIM_FORCE_VIRTUAL 2 /* always virtual */
#define LBSWIM_FORCE_SCATTER 3 /* FP32 storage: the swimmer moves scatter the next step's force (atomics) */

/* lattice fields (LatticeData members, lbswim/liblbswim/Lattice.h:45-49) */
#define LBSWIM_RHO 0
#define LBSWIM_U 1
#define LBSWIM_FORCE 2
#define LBSWIM_FOLD 3
#define LBSWIM_FNEW 4

/* particle fields (SwimmerArray.h:24-28, TracerArray.h:10-13) */
#define LBSWIM_PART_R 0
#define LBSWIM_PART_V 1
#define LBSWIM_PART_N 2    /* swimmers only */
#define LBSWIM_PART_S 3    /* tracers only */
#define LBSWIM_PART_PRNG 4 /* swimmers only: uint64 draw counter per swimmer */

/* kernels that lbswim_lattice_kernel_time() reports */
#define LBSWIM_K_COLLIDE 0 /* DoStep */
#define LBSWIM_K_HYDRO 1   /* DoCalcHydro */
#define LBSWIM_K_ZERO 2    /* DoLatticeZeroForce */
#define LBSWIM_K_ADDF 3    /* DoSwimmerArrayAddForces */
#define LBSWIM_K_SMOVE 4   /* DoSwimmerArrayMove */
#define LBSWIM_K_TMOVE 5   /* DoTracerArrayMove */
#define LBSWIM_K_INIT 6    /* DoInitFromHydro */
#define LBSWIM_K_FUSED 7   /* DoStep + DoCalcHydro in one launch (coupled step) */
#define LBSWIM_K_STORED 8  /* ensembles: stored-force members' force, collide and CalcHydro */
#define LBSWIM_K_COUNT 9

/* Same members, order and types as LBParams (lbswim/liblbswim/LBParams.h:4-30). */
typedef struct lbswim_params {
  double cs2;
  double w[LBSWIM_Q];
  double xi[LBSWIM_Q][LBSWIM_D];
  int ci[LBSWIM_Q][LBSWIM_D];
  double Q[LBSWIM_Q][LBSWIM_D][LBSWIM_D];
  int complement[LBSWIM_Q];
  double norms[LBSWIM_Q];
  double mm[LBSWIM_Q][LBSWIM_Q];
  double mmi[LBSWIM_Q][LBSWIM_Q];
  double tau_s;
  double tau_b;
} lbswim_params;

/* Same as LatticeAddressing (lbswim/liblbswim/Lattice.h:8-12). */
typedef struct lbswim_addressing {
  int size[LBSWIM_D];
  int strides[LBSWIM_D];
  int n;
} lbswim_addressing;

/* Same members, order and layout as CommonParams (lbswim/liblbswim/SwimmerArray.h:8-16). */
typedef struct lbswim_common_params {
  double P;
  double l;
  double alpha;
  double mobility;
  unsigned char translational_advection_off;
  unsigned char rotational_advection_off;
  unsigned long long seed;
} lbswim_common_params;

typedef struct lbswim_lattice lbswim_lattice;
typedef struct lbswim_swimmers lbswim_swimmers;
typedef struct lbswim_tracers lbswim_tracers;

/* ---- library ---------------------------------------------------------- */

/* Message of the last failure on this thread ("" if none). No reference counterpart. */
const char* lbswim_last_error(void);
/* Wait for all enqueued device work (the reference waits inside Step, lb.cu:239). */
int lbswim_synchronize(void);
/* A point in the library's stream: everything enqueued before
 * lbswim_mark_create.  _wait blocks until the device has reached it,
 * _reached tests it, _elapsed times the device between two marks.  Lets
 * a host loop run a bounded distance ahead of the device without draining
 * it.  No reference counterpart. */
typedef struct lbswim_mark lbswim_mark;
int lbswim_mark_create(lbswim_mark** out);
int lbswim_mark_wait(lbswim_mark* m);
int lbswim_mark_reached(lbswim_mark* m, int* reached);
/* Device time (ms) from mark a to mark b, both reached. */
int lbswim_mark_elapsed(lbswim_mark* a, lbswim_mark* b, double* ms);
void lbswim_mark_destroy(lbswim_mark* m);

/* ---- Lattice (lbswim/liblbswim/Lattice.h:53-67; lbswim/_lb.pxd:22-30) ---- */

/* Lattice::Lattice(nx, ny, nz, tau_s, tau_b) (lb.cu:40-63) plus the storage form.
 * Sizes >= 1 with nx*ny*nz < 2^31; tau > 0. Device fields start at zero. */
int lbswim_lattice_create(int nx, int ny, int nz, double tau_s, double tau_b, int storage,
                          lbswim_lattice** out);
/* Lattice::~Lattice (lb.cu:66-67). NULL is ignored. */
void lbswim_lattice_destroy(lbswim_lattice* lat);
/* Lattice::Step: collide + stream, swap fOld/fNew, time_step++ (lb.cu:220-247). */
int lbswim_lattice_step(lbswim_lattice* lat);
/* Lattice::CalcHydro: rho, u from fOld and force (lb.cu:297-312). */
int lbswim_lattice_calc_hydro(lbswim_lattice* lat);
/* Lattice::InitFromHydro: equilibrium fOld from rho, u (lb.cu:373-388). */
int lbswim_lattice_init_from_hydro(lbswim_lattice* lat);
/* Lattice::ZeroForce (lb.cu:413-428). */
int lbswim_lattice_zero_force(lbswim_lattice* lat);
/* Lattice::params.host -- the D3Q15 tables and tau (Lattice.h:61; d3q15.c:2-186). */
int lbswim_lattice_params(lbswim_lattice* lat, const lbswim_params** host);
/* Lattice::addr.host (Lattice.h:62; lb.cu:49-57). */
int lbswim_lattice_addressing(lbswim_lattice* lat, const lbswim_addressing** host);
/* Lattice::time_step read / write (Lattice.h:65; lb.pyx:167-169,240). */
int lbswim_lattice_get_time_step(lbswim_lattice* lat, int* out);
int lbswim_lattice_set_time_step(lbswim_lattice* lat, int ts);
/* Storage form given at creation. No reference counterpart. */
int lbswim_lattice_storage(lbswim_lattice* lat, int* out);
/* LatticeData::{rho,u,force,fOld,fNew} as SharedArray<double>{host,device,size}
 * (Lattice.h:45-49; Shared.h:4-10). *host is the FP64 mirror (stable for the
 * lattice's life), *device the current device buffer in the storage type,
 * *size the element count (n, 3n or 15n). Any output pointer may be NULL.
 * The device distribution buffers hold post-collision values g (streaming is
 * done by pulling); the post-stream fOld the host sees is f_p(x) = g_p(x - c_p),
 * and lbswim_lattice_h2d / lbswim_lattice_d2h apply that shift. */
int lbswim_lattice_field(lbswim_lattice* lat, int field, double** host, void** device, size_t* size);
/* SharedArray::H2D / D2H of one field (Shared.cu:15-26). For fOld/fNew the
 * device side is the buffer currently holding that role, as after the
 * reference's pointer swap (lb.cu:242-244). */
int lbswim_lattice_h2d(lbswim_lattice* lat, int field);
int lbswim_lattice_d2h(lbswim_lattice* lat, int field);

/* ---- SwimmerArray (lbswim/liblbswim/SwimmerArray.h:18-29; SwimmerArray.pxd:16-26) ---- */

/* SwimmerArray::SwimmerArray(n, const CommonParams*) (SwimmerArray.cu:16-25).
 * n >= 0. r, v, n start at zero and every swimmer's draw counter at 0. */
int lbswim_swimmers_create(int n, const lbswim_common_params* p, lbswim_swimmers** out);
void lbswim_swimmers_destroy(lbswim_swimmers* sw);
/* SwimmerArray::num. */
int lbswim_swimmers_num(const lbswim_swimmers* sw, int* out);
/* SwimmerArray::common.host (SwimmerArray.h:25); edit it, then common_h2d. */
int lbswim_swimmers_common(lbswim_swimmers* sw, lbswim_common_params** host);
/* SharedItem<CommonParams>::H2D / D2H (Shared.cu:43-58; swimmers.pyx:21-68). */
int lbswim_swimmers_common_h2d(lbswim_swimmers* sw);
int lbswim_swimmers_common_d2h(lbswim_swimmers* sw);
/* SwimmerArray::AddForces(Lattice*) const (SwimmerArray.cu:114-121). */
int lbswim_swimmers_add_forces(lbswim_swimmers* sw, lbswim_lattice* lat);
/* SwimmerArray::Move(Lattice*) (SwimmerArray.cu:212-220). */
int lbswim_swimmers_move(lbswim_swimmers* sw, lbswim_lattice* lat);
/* SwimmerArray::{r,v,n,prng} as SharedArray{host,device,size}; *elem_size is
 * 8 for all four (double, double, double, uint64 counter).  The host mirror is
 * in the user's order; the device copy is kept in a spatially sorted order
 * that the library changes from time to time (h2d/d2h translate).  Each
 * device pointer stays the same for the array's life (Shared.h:8). */
int lbswim_swimmers_field(lbswim_swimmers* sw, int field, void** host, void** device, size_t* count,
                          size_t* elem_size);
int lbswim_swimmers_h2d(lbswim_swimmers* sw, int field);
int lbswim_swimmers_d2h(lbswim_swimmers* sw, int field);
/* Number of tumbles since creation (diagnostic; no reference counterpart). */
int lbswim_swimmers_tumbles(lbswim_swimmers* sw, unsigned long long* out);
/* *device: int slots[num] on the device, the user index of the swimmer whose
 * data occupy storage slot s of the device fields (component c of field f at
 * device[c*num + s] belongs to swimmer slots[s]).  The pointer is fixed; the
 * contents change when the library re-orders the storage (at a Move after
 * positions were written by h2d, and every 64 moves).  No reference
 * counterpart: the reference keeps the user order on the device. */
int lbswim_swimmers_slots(lbswim_swimmers* sw, const int** device);

/* ---- TracerArray (lbswim/liblbswim/TracerArray.h:6-14; TracerArray.pxd:4-11) ---- */

/* TracerArray::TracerArray(n) (TracerArray.cu:5-10). n >= 0; fields start at zero. */
int lbswim_tracers_create(int n, lbswim_tracers** out);
void lbswim_tracers_destroy(lbswim_tracers* tr);
int lbswim_tracers_num(const lbswim_tracers* tr, int* out);
/* TracerArray::Move(Lattice*) (TracerArray.cu:45-52). */
int lbswim_tracers_move(lbswim_tracers* tr, lbswim_lattice* lat);
/* TracerArray::{r,s,v} (TracerArray.h:10-13); device order as for swimmers. */
int lbswim_tracers_field(lbswim_tracers* tr, int field, void** host, void** device, size_t* count);
int lbswim_tracers_h2d(lbswim_tracers* tr, int field);
int lbswim_tracers_d2h(lbswim_tracers* tr, int field);
/* As lbswim_swimmers_slots, for tracers. No reference counterpart. */
int lbswim_tracers_slots(lbswim_tracers* tr, const int** device);

/* ---- coupled step ------------------------------------------------------ */

/* `nsteps` repetitions of System.Step (lbswim/System.py:9-16): ZeroForce;
 * AddForces of every swimmer array; Step; CalcHydro; Move of every swimmer
 * array, then of every tracer array.
 * Virtual force: with LBSWIM_FORCE_AUTO where the swimmers are sparse (at
 * most about 0.24 (1 + n / 2^18) swimmer ends per 256 of the lattice's n
 * sites; for an ensemble member about 1.9), ZeroForce + AddForces only
 * build sorted per-block point lists and
 * Step and CalcHydro evaluate each site's force from them, bit-identical to
 * the stored force, without writing or reading the force array.  The force
 * array is then written on demand: by lbswim_lattice_d2h(LBSWIM_FORCE),
 * lbswim_lattice_field(LBSWIM_FORCE) with a device pointer requested,
 * lbswim_lattice_sync_force, and before any standalone Step, CalcHydro or
 * AddForces.  A caller that reads the force array on the device by a pointer
 * obtained earlier calls lbswim_lattice_sync_force first. */
int lbswim_system_step(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr,
                       int ntr, int nsteps);
/* lbswim_system_step for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  No reference counterpart.  It calls no hipMalloc, hipFree or
 * allocating library routine: the swimmer-force lists live in a fixed
 * device workspace of 64 MiB per GPU (held from the library's first use on
 * that GPU until the process ends), and the force is written to the
 * lattice's own force array (stored, not virtual).  Results are bit for bit
 * those of lbswim_system_step.  Returns LBSWIM_EINVAL, before enqueueing
 * anything, when the lists of the system need more than the workspace
 * (about 1.4 KB per swimmer plus 12 B per 64 lattice sites: e.g. 256^3 with
 * 32768 swimmers needs 49 MB).  Takes the fused Step+CalcHydro launch only if
 * the lattice already holds its interleaved u buffer (an earlier
 * lbswim_system_step allocated it), else the two launches; tracer moves stay
 * on the library stream.  Every call on a GPU shares that GPU's workspace:
 * calls whose device work overlaps (other streams, concurrent replays of
 * graphs that captured it) race. */
int lbswim_system_step_noalloc(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr,
                               int ntr, int nsteps);
/* LBSWIM_FORCE_STORED, _AUTO (default), _VIRTUAL or (FP32 storage only)
 * _SCATTER. No reference counterpart.
 * Scatter force: where the coupled step runs the fused launch, each step's
 * swimmer moves also add the next step's ZeroForce + AddForces into one of
 * two private force buffers with hardware FP32 atomics -- the reference's own
 * scatter (SwimmerArray.cu:42-112) -- and the fused launch reads it there
 * (clearing the other buffer), so no spreading lists are built.  The order of
 * the atomic additions varies from run to run, and with it the last bit of
 * the FP32 force sums: results agree with the other modes to FP32 rounding
 * (well inside the FP32 form's 1e-5 u error), not bit for bit.  The buffer
 * is kept between calls while the same swimmer arrays are stepped and none
 * of them changed (H2D of r or n, common_h2d, a move outside the call, a
 * device pointer to r or n handed out); otherwise the first step of a call
 * scatters it anew.  The force array itself is written on demand, as for the
 * virtual force.  Lattices stepped by the two launches use the lists. */
int lbswim_lattice_set_force_mode(lbswim_lattice* lat, int mode);
/* The lattice's force mode (lbswim_system_autotune may change it). */
int lbswim_lattice_get_force_mode(lbswim_lattice* lat, int* mode);
/* Writes the force array if the last coupled step left it virtual. No reference counterpart. */
int lbswim_lattice_sync_force(lbswim_lattice* lat);
/* *is_virtual = 1 if the last coupled step left the force virtual (not yet
 * written to the force array). No reference counterpart. */
int lbswim_lattice_force_is_virtual(lbswim_lattice* lat, int* is_virtual);
/* After coupled steps and ensemble steps (below) u lives only in a
 * component-interleaved copy that the particle moves read; the reference-layout u array is written
 * on demand: by lbswim_lattice_d2h(LBSWIM_U), lbswim_lattice_field(LBSWIM_U)
 * with a device pointer requested, lbswim_lattice_init_from_hydro and
 * lbswim_lattice_sync_u.  A caller that reads u on the device by a pointer
 * obtained earlier calls lbswim_lattice_sync_u first.  No reference
 * counterpart. */
int lbswim_lattice_sync_u(lbswim_lattice* lat);
/* *is_virtual = 1 while u is not yet written in the reference layout. */
int lbswim_lattice_u_is_virtual(lbswim_lattice* lat, int* is_virtual);
/* Step fusion in lbswim_system_step: Step and CalcHydro run as one launch
 * that reads the distributions once, with results identical bit for bit to
 * the two launches, on lattices with ny >= rows and 15 n sizeof(storage) <
 * 4 GiB.  LBSWIM_FUSION_AUTO (the default) fuses where the tiles are full
 * (nz % 64 == 0, rows dividing ny), where it is the faster; LBSWIM_FUSION_ON
 * fuses every such lattice (tiles overlap at the edges, lanes idle where nz
 * < 64, and the swimmer force is stored unless nz % 64 == 0);
 * LBSWIM_FUSION_OFF never.  rows: lattice rows per tile (2, 4 or 8; 0
 * chooses); strips: x sweeps per tile column (0 chooses).  No reference
 * counterpart. */
#define LBSWIM_FUSION_OFF 0
#define LBSWIM_FUSION_AUTO 1
#define LBSWIM_FUSION_ON 2
int lbswim_lattice_set_fusion(lbswim_lattice* lat, int enable, int rows, int strips);
/* Chooses how lbswim_system_step runs this system by timing it: the two
 * launches, and the fused launch with tiles of 8, 4 and 2 rows and 1, 2 and
 * 4 times the x strips that give every CU one tile (for an FP32 lattice in
 * LBSWIM_FORCE_AUTO or _SCATTER with swimmers, each fused shape also with
 * the scatter force, which the lattice then keeps if it wins), each over
 * steps_per_trial (+2 warm-up) real coupled steps, then the three fastest
 * three more times each over 2 steps_per_trial (+2) steps -- the system
 * advances by all those steps exactly as lbswim_system_step advances it --
 * and keeps the finalist with the lowest time (set as by
 * lbswim_lattice_set_fusion: LBSWIM_FUSION_OFF, or _ON with those rows and
 * strips).  *rows (0: two launches) and *strips receive
 * the choice; either may be NULL.  No reference counterpart. */
int lbswim_system_autotune(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr,
                           int ntr, int steps_per_trial, int* rows, int* strips);
/* *active = 1 if the coupled step of this lattice runs fused; *rows and
 * *strips (either may be NULL) receive the tile shape it uses. */
int lbswim_lattice_fusion(lbswim_lattice* lat, int* active, int* rows, int* strips);

/* ---- ensembles (no reference counterpart) ------------------------------
 * Many independent systems -- each a lattice with its swimmer and tracer
 * arrays, as in lbswim_system_step -- advanced together: every phase of the
 * coupled step is one launch over all members (descriptor tables on the
 * device), and pairs of steps are replayed from a captured hipGraph.  Each
 * member's result is bit-identical to stepping it alone.  Members stay owned
 * by the caller and must outlive their membership; all members share the
 * ensemble's storage form.  Members may also be used directly between
 * ensemble steps (the ensemble notices and rebuilds its tables). */
typedef struct lbswim_ensemble lbswim_ensemble;
int lbswim_ensemble_create(int storage, lbswim_ensemble** out);
void lbswim_ensemble_destroy(lbswim_ensemble* e);
/* Adds a system; *member receives its slot id (slots of removed members are reused). */
int lbswim_ensemble_add(lbswim_ensemble* e, lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw,
                        lbswim_tracers* const* tr, int ntr, int* member);
/* Takes a member out without waiting for the device: steps already enqueued
 * still complete for it, and later operations on it (copies, a destroy,
 * which waits for the library stream) are ordered after them. */
int lbswim_ensemble_remove(lbswim_ensemble* e, int member);
int lbswim_ensemble_size(lbswim_ensemble* e, int* members);
/* nsteps coupled steps of every member.  The batched steps use the virtual
 * force where the swimmers are sparse; every member's force array holds the
 * last step's force when the call returns. */
int lbswim_ensemble_step(lbswim_ensemble* e, int nsteps);
/* 1 (default): replay captured graphs of two steps; 0: launch eagerly. */
int lbswim_ensemble_set_graph(lbswim_ensemble* e, int on);
/* Workgroup order of the batched launches: 0 member-major, 1 round-robin over
 * members, 2 largest member first. Results do not depend on it. */
int lbswim_ensemble_set_order(lbswim_ensemble* e, int order);
/* LBSWIM_FORCE_STORED, _AUTO (default) or _VIRTUAL for the batched steps. */
int lbswim_ensemble_set_force_mode(lbswim_ensemble* e, int mode);
/* Times each workgroup order with work items of 1, 2 and 4 lattice blocks
 * (256 sites each) over steps_per_trial (+2 warm-up) real steps of the
 * ensemble -- the members advance by those steps -- and keeps the fastest
 * pair, then, with it, the batched fused step's tile shapes
 * (lbswim_ensemble_set_fusion) where some member qualifies, then the number
 * of lanes (lbswim_ensemble_set_lanes); *choice receives the order
 * (lbswim_ensemble_tuning, lbswim_ensemble_fusion and lbswim_ensemble_lanes
 * give the rest). */
int lbswim_ensemble_autotune(lbswim_ensemble* e, int steps_per_trial, int* choice);
/* Lattice blocks of 256 sites per work item (workgroup) of the batched
 * lattice launches, 1..64 (default 1). Results do not depend on it. */
int lbswim_ensemble_set_work(lbswim_ensemble* e, int blocks_per_item);
/* The order and work-item size in use. */
int lbswim_ensemble_tuning(lbswim_ensemble* e, int* order, int* blocks_per_item);
/* Batched fused step: members whose force is virtual, with nz % 64 == 0,
 * rows dividing ny and nx, ny >= 4, advance by one launch of collide/stream
 * + CalcHydro over tiles of `rows` (8 or 4; 0: never) lattice rows x 64
 * sites, their x sweep split into strips of `strip` planes (default 8, 32).
 * Their u is then written component-interleaved and the reference layout is
 * formed on demand (as a single system's fused step).  Results do not
 * depend on it.  lbswim_ensemble_autotune also chooses among off, 8 rows x
 * 16, 32, 64 planes and 4 rows x 32 planes. */
int lbswim_ensemble_set_fusion(lbswim_ensemble* e, int rows, int strip);
int lbswim_ensemble_fusion(lbswim_ensemble* e, int* rows, int* strip);
/* Lattice sites of the members the batched fused step advances. */
int lbswim_ensemble_fused_sites(lbswim_ensemble* e, size_t* fused_sites);
/* Lanes (1..8, default 1): with more than one, the members are split into
 * that many groups of about equal lattice sites (largest lattice first to
 * the lightest group), and each group's batched launches go to a stream of
 * its own, forked from the caller's stream and joined back every step (a
 * captured graph holds them as parallel branches), so that one group's
 * latency-bound list building and moves run beside the others' lattice
 * sweeps.  Results do not depend on it.  lbswim_ensemble_autotune also
 * chooses among 1, 2 and 4 lanes. */
int lbswim_ensemble_set_lanes(lbswim_ensemble* e, int lanes);
int lbswim_ensemble_lanes(lbswim_ensemble* e, int* lanes);
/* Per-phase timing of batched steps with HIP events on the streams the
 * launches go to; while any bit of `mask` (1 << LBSWIM_K_*) is set, steps are
 * launched eagerly instead of from graphs.  Recorded: LBSWIM_K_ADDF (swimmer
 * points and lists), _FUSED (the batched fused step), _COLLIDE and _HYDRO
 * (the other virtual-force members' batched launches), _STORED (the
 * stored-force members' chain), _SMOVE, _TMOVE.  A nonzero mask clears
 * earlier records. */
int lbswim_ensemble_set_timing(lbswim_ensemble* e, unsigned mask);
/* Sum of recorded durations (ms) and launch count of one phase (with lanes,
 * summed over the lanes, whose launches of a step run side by side). */
int lbswim_ensemble_kernel_time(lbswim_ensemble* e, int kernel, double* total_ms, long* count);
/* Lattice sites of the members whose swimmer force is virtual and that the
 * batched _COLLIDE/_HYDRO launches advance, and of the stored-force members
 * (the fused members: lbswim_ensemble_fused_sites). */
int lbswim_ensemble_sites(lbswim_ensemble* e, size_t* virtual_sites, size_t* stored_sites);

/* ---- diagnostics (no reference counterpart) ---------------------------- */

/* FP64 totals of the current distributions: out[0] = total mass, out[1..3] = total momentum. */
int lbswim_lattice_totals(lbswim_lattice* lat, double out[4]);
/* Per-kernel timing with HIP events on the library stream; on = 1 starts
 * recording every kernel (and clears earlier records), 0 stops. */
int lbswim_lattice_set_timing(lbswim_lattice* lat, int on);
/* As set_timing, recording only the kernels whose bit (1 << LBSWIM_K_*) is
 * set: each recorded launch adds an event pair, which costs a few
 * microseconds of device time. */
int lbswim_lattice_set_timing_mask(lbswim_lattice* lat, unsigned mask);
/* Sum of recorded durations (ms) and launch count for one LBSWIM_K_* kernel. */
int lbswim_lattice_kernel_time(lbswim_lattice* lat, int kernel, double* total_ms, long* count);
/* Free and total device memory in bytes (hipMemGetInfo). */
int lbswim_device_memory(size_t* free_bytes, size_t* total_bytes);
/* Copy `count` doubles device->device with a plain streaming kernel `reps` times;
 * *ms is the average per copy (bandwidth probe). */
int lbswim_diag_copy(size_t count, int reps, double* ms);
/* Evaluate the device Peskin delta at x[0..count) (host arrays). */
int lbswim_diag_peskin(const double* x, double* out, int count);
/* Interpolate a host FP64 u[3][nx*ny*nz] at host positions r[3][count] on the device. */
int lbswim_diag_interp(int nx, int ny, int nz, const double* u, const double* r, double* v, int count);

#ifdef __cplusplus
}
#endif

#endif /* LBSWIM_H */
