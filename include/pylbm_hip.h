/*
 * pylbm_hip.h -- C ABI of the MI355X (gfx950) lattice Boltzmann time-stepper.
 *
 * Two kinds of shared objects export these symbols:
 *
 *   libpylbm_hip.so       core runtime: device memory, copies, streams, events,
 *                         and the scheme-independent kernels (periodic ghost
 *                         fill, boundary link lists, FP64 compensated sums).
 *   lbm_gen_<hash>.so     one per scheme, generated from the scheme's moment
 *                         matrix / equilibrium / relaxation rates: the fused
 *                         pull-stream + collide kernel and the standalone
 *                         f2m / m2f / equilibrium kernels, plus a multi-step
 *                         driver.
 *
 * Every entry point returns an int status: 0 = success, > 0 = a hipError_t,
 * < 0 = an argument error; lbm_last_error() returns the message of the last
 * failure on the calling thread.  Pointers named d_* are device pointers,
 * h_* host pointers.  Streams are hipStream_t passed as void* (NULL = the
 * default stream).  No per-step entry point allocates device memory except the
 * link driver when a boundary method needs its two-phase form (it uses a
 * workspace the caller passes in, see lbm_links_apply).
 *
 * The reference routines each entry point replaces (pylbm 0.11.0) are cited
 * as file:line; the Python module pylbm_amd binds these names with ctypes and
 * exposes them under pylbm's routine names (transport / f2m / m2f /
 * equilibrium / one_time_step / bounce_back / anti_bounce_back / neumann*).
 */
#ifndef PYLBM_HIP_H
#define PYLBM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* population storage formats.  LBM_F64 computes in FP64 with full values, as
 * pylbm does.  Every other format computes in FP32 on deviations from the rest
 * state (f_k - w_k rho0, moments m - m_rest), which is the same affine
 * collision in exact arithmetic and avoids the cancellation of O(rho) terms. */
enum lbm_storage {
  LBM_F64 = 0,   /* double                                                   */
  LBM_F32 = 1,   /* float                                                    */
  LBM_F16D = 2,  /* IEEE half of the deviation f_k - w_k rho0               */
  LBM_BF16D = 3, /* bfloat16 of the deviation f_k - w_k rho0                */
  LBM_F32D = 4   /* float of the deviation f_k - w_k rho0                   */
};

#define LBM_MAX_Q 32

/*
 * Device layout of one population array.
 * Spatial axes are canonicalised to (outer, mid, inner): 3D (x, y, z),
 * 2D (1, x, y), 1D (1, 1, x) -- the last reference axis is the contiguous one,
 * as in pylbm's SOA storage (pylbm/storage.py:423-472).  Extents include the
 * halo of width halo[a].  Element (k, o, m, i) lives at
 *   k*pop_stride + (o*n[1] + m)*row_stride + lead + i
 * with pitch (the padded row length) a multiple of 128 bytes and lead chosen so
 * that the first interior element of every row is 128-byte aligned.  Two
 * arrangements: population planes (pop_stride = n0*n1*pitch, row_stride = pitch)
 * and rows interleaved (pop_stride = pitch, row_stride = ns*pitch: the ns rows
 * of one (o, m) line are contiguous).
 */
typedef struct lbm_layout {
  int32_t dim;
  int32_t ns;
  int32_t storage;     /* enum lbm_storage */
  int32_t elem_size;   /* bytes per stored population value */
  int64_t n[3];        /* halo-inclusive extents (outer, mid, inner) */
  int32_t halo[3];
  int32_t _pad;
  int64_t pitch;       /* elements per inner row (padded) */
  int64_t lead;        /* leading pad elements of each row */
  int64_t pop_stride;  /* elements between two populations of one cell */
  int64_t row_stride;  /* elements between two consecutive (o, m) rows */
} lbm_layout;

/* canonical velocity table shared by every kernel of one scheme, and the
 * periodic faces: bit 2a (2a+1) set when the low (high) face of canonical axis a
 * is periodic (pylbm label -1) */
typedef struct lbm_velocities {
  int32_t q;
  int32_t periodic_faces;
  int8_t v[LBM_MAX_Q][3];  /* canonical (outer, mid, inner) components */
} lbm_velocities;

/* ------------------------------------------------------------------ core */
const char* lbm_last_error(void);
int lbm_device_count(int* count);
int lbm_set_device(int device);
int lbm_device_name(int device, char* buf, int len);
int lbm_malloc(void** d_ptr, uint64_t bytes);              /* 256-byte aligned */
int lbm_free(void* d_ptr);
int lbm_memcpy_h2d(void* d_dst, const void* h_src, uint64_t bytes, void* stream);
int lbm_memcpy_d2h(void* h_dst, const void* d_src, uint64_t bytes, void* stream);
int lbm_memcpy_d2d(void* d_dst, const void* d_src, uint64_t bytes, void* stream);
int lbm_memset(void* d_ptr, int value, uint64_t bytes, void* stream);
int lbm_stream_create(void** stream);
int lbm_stream_destroy(void* stream);
int lbm_stream_sync(void* stream);
int lbm_device_sync(void);
int lbm_event_create(void** event);
int lbm_event_destroy(void* event);
int lbm_event_record(void* event, void* stream);
int lbm_event_elapsed_ms(void* start, void* stop, float* ms);
int lbm_event_sync(void* event);                           /* wait for the event on the host */
int lbm_stream_wait_event(void* stream, void* event);     /* later work on stream waits for event */
/* pinned host memory: the target of copies the host reads without draining the stream */
int lbm_host_malloc(void** h_ptr, uint64_t bytes);
int lbm_host_free(void* h_ptr);

/* plain device copy kernel (16-byte vectors): bandwidth reference of the chip
 * and the known-size kernel used to calibrate HBM byte counters */
int lbm_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream);

/*
 * Periodic ghost fill of every halo slot a pull step can read without a link
 * list writing it first: a ghost cell whose first ghost axis (in axis order x,
 * y, z) is a periodic face receives, for every population entering the interior
 * through all of its ghost axes, the value at the fully wrapped interior
 * position.  On those slots this equals pylbm's axis-by-axis wrap of the whole
 * halo (pylbm/storage.py:306-367; Loo.py kernels update_x/y/z 369-420); every
 * other slot a pull reads belongs to a link of a non-periodic label (pylbm flags
 * a corner link with its first axis), which its boundary method overwrites.
 */
int lbm_ghost_fill(void* d_f, const lbm_layout* layout, const lbm_velocities* vel,
                   void* stream);
/* the full axis-by-axis wrap: every population of every ghost cell (used when
 * Neumann-type link lists read ghost slots) */
int lbm_ghost_fill_all(void* d_f, const lbm_layout* layout, const lbm_velocities* vel,
                       void* stream);

/*
 * One boundary link list, f[store[i]] = sign * f[load[i]] + rhs[i], where
 * store/load are element indices into the population array (population offset
 * included) and rhs may be NULL (all zero).  rhs is in the compute type; for the
 * deviation formats (FP32, FP16 and BF16 deviations) it must already include the
 * rest-population shift (lbm_links_prepare_rhs makes it from pylbm's FP64 rhs).
 * Replaces the generated
 * link loops bounce_back (pylbm/boundary.py:430-469), anti_bounce_back
 * (646-684) and neumann / neumannx / neumanny / neumannz (787-895).
 * two_phase != 0 gathers every load into d_work (n values of the compute type)
 * before any store, for lists whose loads and stores overlap (the host has
 * then already folded the sequential read-after-write chains into load/sign/rhs).
 */
int lbm_links_apply(void* d_f, const lbm_layout* layout, const uint32_t* d_store,
                    const uint32_t* d_load, const void* d_rhs, const int8_t* d_sign,
                    double sign, int64_t n, int two_phase, void* d_work, void* stream);

/*
 * One interpolated (Bouzidi) link list,
 *   f[store[i]] = ca[i] * f[load0[i]] + cb[i] * f[load1[i]] + rhs[i],
 * with ca, cb and rhs in the compute type (double for FP64 storage, float
 * otherwise; rhs may be NULL; for the deviation formats rhs includes the
 * rest-population shift of lbm_links_prepare_rhs, which must then be applied
 * even when pylbm's rhs is zero).  Replaces the generated loops Bouzidi_bounce_back
 * (pylbm/boundary.py:506-623: ca = s, cb = 1 - s, loads from a copy of f taken
 * before the loop) and Bouzidi_anti_bounce_back (687-760: ca = -s, cb = 1 - s),
 * where for a wall at distance q the host sets s = 2q, load1 = (ksym, x + 2 v_k)
 * when q < 0.5, and s = 1/(2q), load1 = (k, x + v_k) otherwise.  two_phase != 0
 * gathers every value into d_work (n values of the compute type) before any
 * store: the copy semantics, for lists that load slots they store.
 */
int lbm_links_apply_bouzidi(void* d_f, const lbm_layout* layout, const uint32_t* d_store,
                            const uint32_t* d_load0, const uint32_t* d_load1, const void* d_ca,
                            const void* d_cb, const void* d_rhs, int64_t n, int two_phase, void* d_work,
                            void* stream);

/*
 * The coefficients lbm_links_apply_bouzidi takes, from pylbm's per-link `dist`
 * array of the Bouzidi methods (pylbm/boundary.py:525-535, the `s` it passes as
 * `dist`, FP64): ca = sign * dist, cb = 1 - dist, rounded once to the compute
 * type of `storage` (sign +1: Bouzidi_bounce_back, -1: Bouzidi_anti_bounce_back).
 */
int lbm_links_bouzidi_coefficients(const double* d_dist, double sign, void* d_ca, void* d_cb, int64_t n,
                                   int storage, void* stream);

/*
 * pylbm's FP64 right-hand side of one link list (BoundaryMethod.rhs,
 * pylbm/boundary.py:204-205, set by set_rhs 421-427 / 637-643 / 525-535 and
 * 700-705) in the form lbm_links_apply / lbm_links_apply_bouzidi take it:
 * d_out[i] in the compute type of the layout's storage.  The link kernels of the
 * deviation formats (LBM_F32D, LBM_F16D, LBM_BF16D) combine stored deviations
 * f - w, so for those the rest-population term is added first:
 *   plain list (d_load1 NULL):  rhs[i] + sign w(load0[i]) - w(store[i])
 *   Bouzidi list:               rhs[i] + (sign s w(load0[i]) + (1 - s) w(load1[i]) - w(store[i])),
 *                               s = d_dist[i]
 * with w(j) = h_wshift[population of element index j] (the rest populations
 * w_k rho0 of lbm_params.wshift), evaluated in FP64 and rounded once.  d_rhs may
 * be NULL (all zero; the shift is still applied).  For LBM_F64 / LBM_F32 storage
 * this is a plain conversion (h_wshift and the index arrays may be NULL).  No
 * reference counterpart of its own: the adapter between pylbm's rhs and the
 * deviation storage of this backend.
 */
int lbm_links_prepare_rhs(const lbm_layout* layout, const double* h_wshift, const uint32_t* d_store,
                          const uint32_t* d_load0, const uint32_t* d_load1, const double* d_dist, double sign,
                          const double* d_rhs, void* d_out, int64_t n, void* stream);

/*
 * FP64 compensated (Neumaier) sums of every population over the interior.
 * d_partial needs 2*ns*(nblocks+1) doubles (nblocks from lbm_reduce_blocks),
 * h_sum/h_comp receive ns values each; the conserved sums are
 * sum_k M[c][k] * (h_sum[k] + h_comp[k]).  No reference counterpart: global
 * mass / momentum monitoring of the time-stepper.
 */
int lbm_reduce_blocks(const lbm_layout* layout, int64_t* nblocks);
int lbm_reduce_populations(const void* d_f, const lbm_layout* layout, const double* d_wshift,
                           double* d_partial, double* h_sum, double* h_comp, void* stream);

/*
 * lbm_reduce_populations without the d_partial workspace and without any
 * allocation: the partial sums go to a fixed scratch array of the library on the
 * current device (2 * LBM_MAX_Q * 1025 doubles, 512.5 KiB, held by every GPU the
 * library runs on until the process ends; lbm_set_device looks up its address).
 * Same kernels, same grid and same summation order, so the same h_sum / h_comp
 * bit for bit.  Layouts whose reduction would need more scratch are refused with
 * an error (none does within LBM_MAX_Q populations).  Calls that overlap on one
 * GPU (several host threads or streams at once) share the scratch array and race;
 * each call writes every partial it reads before reading it, so nothing needs
 * clearing between calls.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under graph capture, tight monitoring
 * loops, harnesses that forbid hipMalloc).
 */
int lbm_reduce_populations_noalloc(const void* d_f, const lbm_layout* layout, const double* d_wshift,
                                   double* h_sum, double* h_comp, void* stream);

/* ---------------------------------------------------- generated, per scheme */
/*
 * Run-time parameters of the generated kernels.  omega[i] = s_i and
 * one_minus[i] = 1 - s_i for every relaxed moment (pylbm/algorithm/base.py:404-428);
 * wshift[k] = w_k rho0 = (M^-1 eq(cref))_k is the rest population that the
 * FP32-compute formats subtract before the collision.
 */
typedef struct lbm_params {
  double omega[LBM_MAX_Q];
  double one_minus[LBM_MAX_Q];
  double wshift[LBM_MAX_Q];
  double cref[LBM_MAX_Q];   /* reference conserved moments (rest state) of the deviation formats */
  double dt;
  double t;
  /* Link lists folded into the step's pulls.  bbmask: NULL, or one word per
   * interior cell (outer, mid, inner order); bit k set means the pull of
   * population k at this cell does not read f[k, x - v_k] but computes the value
   * the step's link lists (pylbm/simulation.py:385-390; bounce-back,
   * anti-bounce-back, Neumann, pylbm/boundary.py:411-427, 637-643, 777-890) would
   * have left there, with the link kernel's arithmetic: a copy of f[ksym, x]
   * unless bit 31 of the mask word is set (ns <= 31) and the cell's bbtab_bits
   * word has bit k set; then entry
   * e = bbtab_base[cell] + (number of lower set bits) gives
   * sign[e] * f[bbtab_src[e]] + rhs[e] (sign / rhs in the compute type: double for
   * FP64 storage, float otherwise; src a flat element index).  The host folds
   * only what this reproduces exactly and leaves folded lists out of the step. */
  const uint32_t* bbmask;
  const uint32_t* bbtab_bits;
  const uint32_t* bbtab_base;
  const uint32_t* bbtab_src;
  const void* bbtab_sign;
  const void* bbtab_rhs;
  /* NULL, or per table entry its source relative to the pulling cell, for the
   * temporally blocked run: bits 0-7 the source population, bits 8-15 and 16-23
   * the mid / inner offsets + 128 (within one cell; 2D schemes) */
  const uint32_t* bbtab_rel;
  /* NULL, or per table entry a second source for the interpolated (Bouzidi)
   * links of pylbm/boundary.py:472-760: where bbtab_src2[e] is not 0xffffffff the
   * entry is sign[e] * f[src[e]] + c2[e] * f[src2[e]] + rhs[e] (the arithmetic of
   * lbm_links_apply_bouzidi, ca = sign, cb = c2, each product and sum rounded
   * on its own).  bbtab_rel2: the relative form of src2 (as bbtab_rel) with bit
   * 31 set on the entries that have a second source; the others repeat their
   * bbtab_rel word.  The temporally blocked run needs bbtab_rel2 when bbtab_src2
   * is set. */
  const uint32_t* bbtab_src2;
  const void* bbtab_c2;
  const uint32_t* bbtab_rel2;
  /* cell-centre coordinates for source terms that depend on space (pylbm's X, Y,
   * Z): reference axis d of the cell at halo-inclusive index i is at
   * origin[d] + dx * i (the coordinates of pylbm's Domain.coords_halo) */
  double origin[3];
  double dx;
  /* workgroup order of the step: 1 = each of the 8 XCDs takes a contiguous range
   * of rows (their shared neighbour rows then meet in one L2), 0 = the hardware's
   * round-robin order, -1 = the default (1 on the halo path, 0 on the wrap path) */
  int32_t xcd_remap;
  /* cells per thread of the step along the contiguous axis: 0 = the default (4
   * for the 2- and 4-byte formats where that leaves >= 2^20 threads, else 1),
   * 1 = one cell per thread, 4 = four (the contiguous interior extent must be a
   * multiple of 4); every choice gives the same bits */
  int32_t vec_cells;
} lbm_params;

/* information on the generated module (ns, nconsm, dim) */
int lbm_gen_info(int32_t* ns, int32_t* nconsm, int32_t* dim);

/*
 * One fused pull step on the interior: for every interior cell x,
 * m = M f(x - v), [source dt/2], relax, [source dt/2], fnew(x) = M^-1 m.
 * Replaces the generated one_time_step loop of
 * pylbm/algorithm/pull.py:11-59 (registered by pylbm/algorithm/base.py:551-628).
 * periodic_fold != 0 reads pulled values through the periodic wrap of every
 * axis instead of the halo: valid for a periodic box, and for any box whose link
 * lists are all folded into bbmask / bbtab (the ghost slots the pulls would read
 * then hold the wrapped interior values, pylbm/storage.py:306-367, or are
 * replaced by the folded link values), with the table sources in the interior.
 */
int lbm_gen_step(const void* d_f, void* d_fnew, const lbm_layout* layout,
                 const lbm_params* params, int periodic_fold, void* stream);

/* f2m on the whole halo-inclusive domain into FP64 moments d_m (layout of m
 * is `layout` with storage LBM_F64). pylbm/algorithm/base.py:298-344 */
int lbm_gen_f2m(const void* d_f, const lbm_layout* flayout, double* d_m,
                const lbm_layout* mlayout, const lbm_params* params, void* stream);
/* m2f on the whole domain from FP64 moments. pylbm/algorithm/base.py:346-379 */
int lbm_gen_m2f(const double* d_m, const lbm_layout* mlayout, void* d_f,
                const lbm_layout* flayout, const lbm_params* params, void* stream);
/* equilibrium in place on FP64 moments. pylbm/algorithm/base.py:381-402 */
int lbm_gen_equilibrium(double* d_m, const lbm_layout* mlayout, void* stream);
/* lbm_gen_equilibrium then lbm_gen_m2f in one pass (the moments stay in registers;
 * only the conserved rows of d_m are read, none is written): the moment
 * initialisation of pylbm/simulation.py:314-316 */
int lbm_gen_eq_m2f(const double* d_m, const lbm_layout* mlayout, void* d_f,
                   const lbm_layout* flayout, const lbm_params* params, void* stream);
/* pylbm's split routines, as Simulation.transport / relaxation / source_term call
 * them (pylbm/simulation.py:322-342):
 * transport: d_fnew[k, x] = d_f[k, x - v_k] on the interior (a copy of the stored
 * words; the ghost slots it reads are the caller's). pylbm/algorithm/base.py:273-296 */
int lbm_gen_transport(const void* d_f, void* d_fnew, const lbm_layout* flayout,
                      const lbm_params* params, void* stream);
/* relaxation in place on FP64 moments over the halo-inclusive domain,
 * m_i <- (1 - s_i) m_i + s_i eq_i (params->omega / one_minus). pylbm/algorithm/base.py:404-437 */
int lbm_gen_relaxation(double* d_m, const lbm_layout* mlayout, const lbm_params* params, void* stream);
/* source terms in place on FP64 moments over the interior: one explicit Euler half
 * step m_c <- m_c + dt/2 S(m, X, Y, Z, t) (params->dt, t, origin, dx).
 * pylbm/algorithm/base.py:439-504, pylbm/algorithm/ode.py:11-16 */
int lbm_gen_source_term(double* d_m, const lbm_layout* mlayout, const lbm_params* params, void* stream);

/*
 * nsteps complete time steps (ghost fill, link lists in order, fused step,
 * swap) without returning to the host: pylbm/simulation.py:392-420 repeated.
 * d_f / d_fnew are swapped nsteps times; on return *parity = nsteps % 2 tells
 * which buffer holds the newest state (0: d_f).  links: nmeth descriptors.
 */
typedef struct lbm_link_set {
  const uint32_t* d_store;
  const uint32_t* d_load;
  const void* d_rhs;
  const int8_t* d_sign;
  double sign;
  int64_t n;
  int32_t two_phase;
  /* 1: no slot this list loads or stores is stored by the lists of the group it
   * joins (the previous list and those it joined), and it loads no slot those
   * lists store: it may run in the same launch as them (set by the host) */
  int32_t merge_prev;
  void* d_work;
  /* Interpolated (Bouzidi) list when d_load1 is not NULL:
   * f[store] = d_ca[i] f[load] + d_cb[i] f[load1] + rhs (compute-type arrays, see
   * lbm_links_apply_bouzidi; d_sign unused); merge_prev as for the other lists */
  const uint32_t* d_load1;
  const void* d_ca;
  const void* d_cb;
} lbm_link_set;

/*
 * The link lists of one time step in order (pylbm/simulation.py:385-390), in as
 * few launches as their merge_prev flags allow: a group of up to 8 consecutive
 * one-phase lists joined by merge_prev runs as one kernel, with each link
 * computed as lbm_links_apply (or, d_load1 set, lbm_links_apply_bouzidi) computes it.
 */
int lbm_links_apply_sets(void* d_f, const lbm_layout* layout, const lbm_link_set* sets, int nsets,
                         void* stream);

/*
 * nsteps time steps of a 2D scheme on the wrap path (a periodic box, or every link
 * list folded into bbmask / bbtab with bbtab_rel set), steps_per_launch (<= 16) of
 * them per launch: each workgroup loads a tile_m x tile_i tile of the interior with
 * a halo of steps_per_launch cells into LDS and advances it there, with the
 * per-cell arithmetic of lbm_gen_step (the same bits); the remaining
 * nsteps % steps_per_launch steps run as lbm_gen_step launches.  This is
 * pylbm/simulation.py:392-420 repeated, for grids whose step is too small to fill
 * the GPU.  *parity as for lbm_gen_run; params->t advances by nsteps dt.
 */
int lbm_gen_run_blocked(void* d_f, void* d_fnew, const lbm_layout* layout, lbm_params* params,
                        int64_t nsteps, int steps_per_launch, int tile_m, int tile_i, int* parity,
                        void* stream);

int lbm_gen_run(void* d_f, void* d_fnew, const lbm_layout* layout, const lbm_velocities* vel,
                lbm_params* params, const lbm_link_set* links, int nmeth, int64_t nsteps,
                int ghost, int periodic_fold, int* parity, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYLBM_HIP_H */
