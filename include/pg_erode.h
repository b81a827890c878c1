/*
 * liberode_hip -- the particle hydraulic erosion droplet loop on the
 * equi-angular cube-sphere, as HIP kernels for MI355X (gfx950).
 *
 * C ABI.  Every pointer argument named like a state array is a DEVICE
 * pointer (e.g. a PyTorch-ROCm tensor's data_ptr()); scalars are by value;
 * parameter structs are host pointers.  Arrays are C-contiguous with the
 * reference's shapes: extended fields are (F, NE, NE[, C]) with
 * NE = N + 2H, flattened as ((f * NE + i) * NE + j) * C + c.
 * The caller owns every buffer (as in the reference, bake/globe/erosion/
 * maps.py:933-944); no entry point allocates device memory.
 * Every call returns 0 on success, non-zero on error; pg_last_error()
 * returns the message of the last failure on the calling thread.  No C++
 * exception crosses the ABI.  Work is enqueued on `stream` (a hipStream_t;
 * NULL = the default stream) and is asynchronous unless stated otherwise.
 *
 * Modes.  The exact (check) mode reproduces the reference's algorithm bit
 * for bit: float64 droplet state, the packed float32 record, chunked trace
 * against frozen maps (pg_trace_particles), then the chunk's change lists
 * applied with the serial apply's results -- computed in parallel by
 * pg_apply_changes_ordered (or serially by pg_apply_changes), the whole
 * chunk loop in one call by pg_check_chunks.  The fast mode (pg_fast_*)
 * traces every droplet of a chunk in float32 and accumulates its effects as
 * int64 fixed point with atomics: deterministic, exactly mass-conserving,
 * with order-independent caps -- a different algorithm, far from the
 * reference's results at the earth globe's size.
 */
#ifndef PG_ERODE_H
#define PG_ERODE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* pg_stream; /* hipStream_t */

/* Scalar parameters of trace_particles, same names, order and meaning as
 * its arguments (bake/globe/erosion/particle.py:292-318). */
typedef struct pg_trace_params {
    double dt, density, friction, deposition_rate, evap_rate, k_mom, mom_scale, k_disc, disc_saturation,
        disc_exponent, slope_gain, slope_sat, erodibility, cover_depth, min_volume;
    int64_t max_steps;
    double max_erode, conc_scale;
    int64_t pit_steps;
    int64_t deposit_on_exit;
    double ocean_rate;
    int64_t ocean_steps;
    double fan_room, dep_floor, lake_trap, lake_park, lateral_rate;
} pg_trace_params;

/* Message of the last failed call on this thread ("" if none). */
const char* pg_last_error(void);
/* ABI version of this library: 9 (bumped whenever a struct layout or a
 * signature changes; 2 added pg_fast_params.erode_cap, 3 its tape, 6 the
 * brush, 7 the deterministic adjoint's workspace and halo table, 8 the
 * adjoint's tape in pieces, 9 the fast mode's bank cuts and their stat). */
int pg_abi_version(void);

/* ---- geometry, host side (bake/globe/cubesphere.py:70-220) ------------- */
/* The library's own scalar geometry run on the host CPU: the same code the
 * kernels run, exposed so it can be checked without a GPU.
 * Replace to_sphere (cubesphere.py:70), from_sphere (:96), jacobian (:130)
 * and transfer_velocity (:204); pg_geom_tan/atan/sincos are the
 * double-double functions behind them (math.tan / math.atan / math.cos,
 * math.sin as numba calls them, particle.py:400-401). */
int pg_geom_to_sphere(int face, double u, double v, double* xyz3);
int pg_geom_from_sphere(double x, double y, double z, int* face, double* uv2);
int pg_geom_jacobian(int face, double u, double v, double* j6);
int pg_geom_transfer_velocity(int face, double u, double v, double a, double b, int* face_out, double* uvab4);
/* The fast mode's face transfer: transfer_velocity in single precision
 * (float tan / atan; no counterpart in the reference, whose droplets are
 * float64), on the host. */
int pg_geom_transfer_velocity_f32(int face, float u, float v, float a, float b, int* face_out, float* uvab4);
double pg_geom_tan(double x);
double pg_geom_atan(double x);
void pg_geom_sincos(double x, double* s, double* c);
/* The same functions evaluated on the GPU over DEVICE arrays of n inputs:
 * op 0 tan (in n, out n), 1 atan (in n, out n), 2 to_sphere (in n x (face,
 * u, v), out n x 3), 3 from_sphere (in n x 3, out n x (u, v), iout face),
 * 4 transfer_velocity (in n x (face, u, v, a, b), out n x (u, v, a, b),
 * iout face), 5 sin/cos (in n, out n x (sin, cos)), 6 the fast mode's
 * single-precision transfer (as 4, computed in float32). */
int pg_geom_eval(int op, int64_t n, const double* in, double* out, int* iout, pg_stream stream);

/* ---- packed samples (particle.py:164-191, pack_samples) ---------------- */
/* samp (F, NE, NE, 16) float32 from the state arrays.  route may be NULL
 * when use_route == 0; lake_flag may be NULL (all zero). */
int pg_pack_samples(float* samp, int64_t F, int64_t NE, const double* height, const double* sediment,
                    const double* route, int use_route, const double* discharge, const double* momentum,
                    const float* evap, const float* hardness, const float* metric, const float* metric_inv,
                    const double* base, const uint8_t* lake_flag, pg_stream stream);

/* ---- check mode: trace (particle.py:275-818, trace_particles) ---------- */
/* One chunk of P droplets on frozen samples, float64 droplet state.  Fills
 * the change list exactly like the reference: P * cap entries with
 * cap = pg_change_list_cap(max_steps, lateral_rate > 0).  sp_dir is read
 * only where sp_sed > 0.  roots may be NULL when use_roots == 0. */
int pg_trace_particles(int64_t P, const int64_t* sp_face, const double* sp_x, const double* sp_y,
                       const double* sp_sed, const double* sp_dir, double volume0, const float* samp, int use_route,
                       const uint8_t* mask, int64_t N, int64_t H, int spherical, const pg_trace_params* prm,
                       const float* roots, int use_roots, int32_t* cl_cell, double* cl_delta, float* cl_vol,
                       float* cl_mom, int64_t* cl_count, int64_t* cl_lo, int64_t* cl_hi, int8_t* sp_death,
                       pg_stream stream);

/* Droplets per wave of pg_trace_particles (and of the chunk loop): a power
 * of two in [1, 64] (default 64); the same results at every setting (each
 * lane traces its own droplet).  Library-wide state, not thread-safe.  No
 * counterpart in the reference (a launch shape). */
int pg_check_set_lanes(int lanes);

/* particle.py:258-272, change_list_cap */
int64_t pg_change_list_cap(int64_t max_steps, int lateral);

/* ---- check mode: serial apply (particle.py:821-1042, apply_changes) ---- */
/* Applies the change list in droplet order against the live terrain
 * (height + sediment + acc), updating acc, pending, samp (S_SURF, S_SED,
 * S_RFLAG), the discharge/momentum tracks and the lake loads.  Writes
 * (clamped entries, mass to pending, deficit lost, lost offshore) to the
 * DEVICE array out4 (float64[4]).  erode_cap may be NULL when
 * use_erode_cap == 0; lake_load / lake_id / lake_room may be NULL when the
 * list holds no lake entries (volume -3). */
int pg_apply_changes(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                     const int64_t* cl_count, int64_t P, int64_t cap, const double* height, const double* sediment,
                     double* acc, double* pending, float* samp, double* disch_track, double* mom_track,
                     const uint8_t* mask, double iter_erode, double iter_deposit, double fan_slope, int use_route,
                     double dep_floor, double offshore_writeoff, const double* erode_cap, int use_erode_cap,
                     double* lake_load, const int32_t* lake_id, double* lake_room, double* out4, pg_stream stream);

/* ---- check mode: the same apply in parallel, bit for bit ------------- */
/* Replaces apply_changes (particle.py:821-1042) like pg_apply_changes, with
 * the same results to the bit, computed in parallel: the serial order's
 * values are the unique solution of "each entry's amount from the live
 * values it reads; each read value from the earlier amounts of its cell",
 * solved by alternating droplet walks and per-cell walks in serial order
 * until a sweep changes nothing (at most max_sweeps; 10-170 on the earth
 * globe, thousands on clamp-dominated windows).  The first launched_sweeps sweeps are enqueued as plain kernel
 * launches that return at once after convergence; a chunk needing more
 * finishes in one cooperative launch.  n_cells = F * NE * NE.  status:
 * DEVICE int32[4], zeroed by the caller once per iteration; the call ORs
 * bit 1 into status[0] when the list holds lake parks (volume -3, not
 * supported: the chunk is then not applied), bit 2 when max_sweeps did not
 * converge (not applied); status[1] = most sweeps of a chunk, status[2] +=
 * this chunk's sweeps.  work: caller-allocated DEVICE scratch of
 * pg_ordered_workspace_bytes(P, cap, max_sweeps) bytes.  No host
 * synchronisation. */
int64_t pg_ordered_workspace_bytes(int64_t P, int64_t cap, int max_sweeps);
int pg_apply_changes_ordered(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol, const float* cl_mom,
                             const int64_t* cl_count, int64_t P, int64_t cap, const double* height,
                             const double* sediment, double* acc, double* pending, float* samp, double* disch_track,
                             double* mom_track, const uint8_t* mask, int64_t n_cells, double iter_erode,
                             double iter_deposit, double fan_slope, int use_route, double dep_floor,
                             double offshore_writeoff, const double* erode_cap, int use_erode_cap, double* out4,
                             int32_t* status, int max_sweeps, int launched_sweeps, void* work, int64_t work_bytes,
                             pg_stream stream);

/* No counterpart in the reference: pg_apply_changes_ordered for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results, no
 * workspace: the scratch is a file-scope device array of
 * pg_ordered_scratch_bytes() bytes (96 MiB: chunks of up to 2048 droplets
 * with lifetimes up to 128 and max_sweeps up to 256 fit), looked up once per
 * device and held by every GPU the library runs on until the process ends.
 * Sizes beyond it return an error (nothing is launched).  Calls that overlap
 * on one GPU share the array and race.  No cooperative launch: every sweep is
 * a plain launch, max_sweeps (<= 4096) of them at most.  Calls no hipMalloc
 * or allocating routine. */
int64_t pg_ordered_scratch_bytes(void);
int pg_apply_changes_ordered_noalloc(const int32_t* cl_cell, const double* cl_delta, const float* cl_vol,
                                     const float* cl_mom, const int64_t* cl_count, int64_t P, int64_t cap,
                                     const double* height, const double* sediment, double* acc, double* pending,
                                     float* samp, double* disch_track, double* mom_track, const uint8_t* mask,
                                     int64_t n_cells, double iter_erode, double iter_deposit, double fan_slope,
                                     int use_route, double dep_floor, double offshore_writeoff,
                                     const double* erode_cap, int use_erode_cap, double* out4, int32_t* status,
                                     int max_sweeps, pg_stream stream);

/* The check mode's chunk loop (bake/globe/erosion/maps.py:953-1029): for each
 * chunk of `chunk` droplets of the P spawned, pg_trace_particles on the frozen
 * record, pg_apply_changes_ordered, pg_chunk_stats -- the same results as
 * calling them chunk by chunk, without a host round trip per chunk.  The
 * change-list buffers hold chunk * cap entries; outs (DEVICE float64
 * [n_chunks * 4]) receives each chunk's 4-tuple in chunk order, summed in
 * droplet order from drop_out (DEVICE float64 [P * 4] scratch: each
 * droplet's clamped entries, stockpile, deficit and offshore loss).  moves
 * (DEVICE int64, may be NULL) += the steps traced (entries of volume >= 0);
 * timing != 0 records HIP events around every trace launch.  A chunk whose
 * lists hold lake parks (volume -3), or that does not converge within
 * max_sweeps, is applied by the serial apply instead (pg_apply_changes'
 * kernel, enqueued behind a device-side test: no host round trip; the same
 * results); status[3] counts those chunks, and status[0] stays 0.
 * lake_load / lake_id / lake_room as in pg_apply_changes (may be NULL when no
 * list holds lake entries). */
int pg_check_chunks(int64_t P, int64_t chunk, const int64_t* sp_face, const double* sp_x, const double* sp_y,
                    const double* sp_sed, const double* sp_dir, double volume0, float* samp, int use_route,
                    const uint8_t* mask, int64_t N, int64_t H, int spherical, const pg_trace_params* prm,
                    const float* roots, int use_roots, int32_t* cl_cell, double* cl_delta, float* cl_vol,
                    float* cl_mom, int64_t* cl_count, int64_t* cl_lo, int64_t* cl_hi, int8_t* sp_death,
                    const double* height, const double* sediment, double* acc, double* pending,
                    double* disch_track, double* mom_track, double iter_erode, double iter_deposit,
                    double fan_slope, double dep_floor, double offshore_writeoff, const double* erode_cap,
                    int use_erode_cap, double* lake_load, const int32_t* lake_id, double* lake_room,
                    double* outs, int64_t* counts7, int32_t* status, int max_sweeps,
                    int launched_sweeps, void* work, int64_t work_bytes, double* drop_out, int64_t* moves,
                    int timing, pg_stream stream);
/* Summed duration (ms) and count of the trace launches pg_check_chunks timed
 * with timing != 0 since the last call (HIP events on the launch stream);
 * synchronises on the last one. */
int pg_check_trace_time(double* ms_total, int64_t* launches);

/* Adds a chunk's entry count and death histogram to the DEVICE array
 * counts7 (int64[7]: entries, then the six DEATH_* causes), the statistics
 * run_iteration gathers per chunk (bake/globe/erosion/maps.py:1028-1029). */
int pg_chunk_stats(const int64_t* cl_count, const int8_t* sp_death, int64_t P, int64_t* counts7, pg_stream stream);

/* ---- per-cell passes ---------------------------------------------------- */
/* particle.py:1079-1101, fold_changes (n = F * NE * NE) */
int pg_fold_changes(int64_t n, double* height, double* sediment, double* acc, const uint8_t* mask, pg_stream stream);
/* particle.py:1230-1242, ema_update */
int pg_ema_update(int64_t n, double* discharge, double* momentum, double* disch_track, double* mom_track,
                  const uint8_t* mask, double ema, pg_stream stream);
/* bake/globe/erosion/maps.py:1132-1134, apply_uplift: height += min(u, cap) - mean
 * on active cells (cap ignored when use_cap == 0; mean from the caller). */
int pg_apply_uplift(int64_t n, double* height, const double* uplift, const uint8_t* mask, double cap, int use_cap,
                    double mean, pg_stream stream);

/* ---- spawning (particle.py:1245-1260, spawn_cells) --------------------- */
/* out[p] = index of the first cdf entry > u[p] * cdf[M-1] (capped at M-1);
 * u sorted ascending. */
int pg_spawn_cells(const double* cdf, int64_t M, const double* u, int64_t P, int64_t* out, pg_stream stream);

/* ---- halo exchange (bake/globe/field.py:104-124 with the tables of
 * bake/globe/cubesphere.py:540-662) -------------------------------------- */
/* data (6*NE*NE, ncomp) float64.  Scalar (ncomp 1, rot NULL):
 * data[dst[k]] = sum_s w[k,s] data[src[k,s]].  Vector (ncomp 2, rot given):
 * data[dst[k]] = sum_s w[k,s] rot[k,s] . data[src[k,s]].  Sources are
 * interior cells, destinations halo cells. */
int pg_exchange_halos(double* data, int64_t ncomp, int64_t K, int64_t S, const int64_t* dst, const int64_t* src,
                      const double* w, const double* rot, pg_stream stream);

/* Adjoint of the scalar exchange (no counterpart in the reference, which has
 * no adjoint): adj[src[k,s]] += w[k,s] adj[dst[k]], then adj[dst[k]] = 0 --
 * the halo entries of a gradient folded onto the interior cells they are
 * interpolated from.  Deterministic: the inverse table lists, per source
 * cell src_cell[i], the entries k * S + s that read it in increasing order
 * (src_entry[src_start[i] .. src_start[i + 1])); each cell's contributions
 * are summed in that order and added once.  Sources must not be
 * destinations.  All arrays DEVICE. */
int pg_halo_scalar_adjoint(double* adj, int64_t K, int64_t S, const int64_t* dst, const double* w, int64_t n_src,
                           const int64_t* src_cell, const int64_t* src_start, const int64_t* src_entry,
                           pg_stream stream);

/* ---- datum hold (bake/globe/erosion/maps.py:1446-1508, hold_datum) ----- */
/* Exact k-th smallest (0-based) of height + sediment over the interior
 * (F faces, N x N, halo H), written to the DEVICE double *q_out.
 * work: DEVICE scratch of pg_select_work_bytes() bytes. */
int64_t pg_select_work_bytes(void);
int pg_select_kth_surface(const double* height, const double* sediment, int64_t F, int64_t N, int64_t H, int64_t k,
                          double* q_out, void* work, pg_stream stream);
/* height -= q; route -= q (when route != NULL); base -= q where base < 0;
 * q read from the DEVICE double *q (n = F * NE * NE). */
int pg_datum_shift(int64_t n, double* height, double* route, double* base, const double* q, pg_stream stream);

/* ---- mass wasting (bake/globe/erosion/maps.py:1050-1085, thermal_erosion) */
/* talus[c] = soft + (hard - soft) * hardness[c] in float32, widened
 * (maps.py:1062). */
int pg_talus(int64_t n, const float* hardness, double soft, double hard, double* talus, pg_stream stream);
/* particle.py:1128-1160, thermal_pass_a: per-cell cap scale (and total
 * outflow into out_total unless it is NULL).  F faces of NE x NE. */
int pg_thermal_pass_a(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                      const double* talus, const uint8_t* mask, double rate, double thermal_max, double* out_total,
                      double* scale, pg_stream stream);
/* particle.py:1163-1187, thermal_pass_r: receiver scale of submerged cells. */
int pg_thermal_pass_r(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                      const double* talus, const uint8_t* mask, double rate, const double* scale, double* rscale,
                      double dep_floor, pg_stream stream);
/* particle.py:1190-1227, thermal_pass_b: the gather into new_height /
 * new_sediment (which must not alias height / sediment). */
int pg_thermal_pass_b(int64_t F, int64_t NE, const double* height, const double* sediment, const float* metric,
                      const double* talus, const uint8_t* mask, double rate, const double* scale, const double* rscale,
                      double* new_height, double* new_sediment, pg_stream stream);
/* maps.py:1073-1085, _mass_wasting_pass: the three passes and the copy back
 * into height / sediment.  work: DEVICE float64[4 * F * NE * NE]. */
int pg_mass_wasting_pass(int64_t F, int64_t NE, double* height, double* sediment, const float* metric,
                         const double* talus, const uint8_t* mask, double rate, double thermal_max, double dep_floor,
                         double* work, pg_stream stream);
/* maps.py:1067: out = FROZEN where height + sediment < base, else mask. */
int pg_creep_mask(int64_t n, const double* height, const double* sediment, const double* base, const uint8_t* mask,
                  uint8_t* out, pg_stream stream);

/* ---- slope limit (particle.py:1045-1076, slope_erode_cap) -------------- */
int pg_slope_erode_cap(int64_t F, int64_t NE, const double* height, const double* sediment, const uint8_t* mask,
                       double coef, double exit_slope, double* out, pg_stream stream);

/* ---- isostasy (maps.py:1325-1410) and its smoother (field.py:179-188,
 * 372-404) -------------------------------------------------------------- */
/* The flux-form Laplace-Beltrami operator of data (F x NE x NE, halos
 * exchanged), lap = float32(raw / cell_size_sq) as FaceField.laplacian:
 * mode 0: out = raw (the operator before the division; 0 on the outer ring);
 * mode 1: out = lap * c2;  mode 2: out = data + kappa * lap * c2 (one
 * smoothing step).  metric_inv float32 (F, NE, NE, 3), sqrt_det float32. */
int pg_laplacian(int64_t F, int64_t NE, const double* data, const float* metric_inv, const float* sqrt_det,
                 double cell_size_sq, int32_t mode, double kappa, double c2, double* out, pg_stream stream);
/* Over the interior cells (mask NULL: all; else mask == ACTIVE only):
 * out4 = {sum of v (what 0) / v^2 (1) / |v| (2), count, max v, min v}, a
 * DEVICE double[4].  work: DEVICE scratch of pg_interior_sums_work_bytes(). */
int64_t pg_interior_sums_work_bytes(void);
int pg_interior_sums(const double* data, const uint8_t* mask, int64_t F, int64_t N, int64_t H, int32_t what, double* work,
                     double* out4, pg_stream stream);
/* maps.py:1405-1406: height += data - mean on active cells, with mean =
 * sums[0] / sums[1] (0 when sums[1] == 0) read from the DEVICE sums. */
int pg_rebound(int64_t n, double* height, const double* data, const uint8_t* mask, const double* sums, pg_stream stream);
/* out = height + sediment (maps.py:1574). */
int pg_surface(int64_t n, const double* height, const double* sediment, double* out, pg_stream stream);
/* iso_acc += (height + sediment) - s0 (maps.py:1586). */
int pg_iso_accumulate(int64_t n, double* iso_acc, const double* height, const double* sediment, const double* s0,
                      pg_stream stream);
/* out = a * x (a == 0: zeros, x may be NULL); out = x / v. */
int pg_scale(int64_t n, const double* x, double a, double* out, pg_stream stream);
int pg_divide(int64_t n, const double* x, double v, double* out, pg_stream stream);

/* ---- routing surface (bake/globe/erosion/route.py:30-117,
 * priority_flood_eps) ----------------------------------------------------- */
/* The flood's epsilon-filled surface as the fixed point of
 * W(n) = max(s(n), min over in-neighbours W + eps), relaxed in place.
 * cls (uint8, F x NE x NE): 0 never assigned (outside / outer ring / a
 * cross-face halo cell), 1 static seed (an interior cell next to the outside
 * or the window edge), 2 assignable, 3 assignable with extra in-edges listed
 * in spec_cell (sorted, n_spec) / spec_ptr (n_spec + 1) / spec_src.
 * pg_flood_init: s = height + sediment, W = s on seeds (also s < 0 and
 * s <= base on interior cells), +inf elsewhere; *flag = 0.
 * pg_flood_sweeps: `sweeps` relaxation sweeps; *flag (DEVICE int32) is 1 when
 * the last of them changed a value (0: converged).
 * pg_flood_finish: route = W where reached, s elsewhere. */
int pg_flood_init(int64_t F, int64_t NE, int64_t H, int64_t N, const double* height, const double* sediment,
                  const double* base, const uint8_t* cls, double* s_out, double* W, int32_t* flag, pg_stream stream);
int pg_flood_sweeps(int64_t F, int64_t NE, const double* s, const uint8_t* cls, double* W, double eps,
                    const int64_t* spec_cell, const int32_t* spec_ptr, const int64_t* spec_src, int64_t n_spec,
                    int32_t sweeps, int32_t* flag, pg_stream stream);
int pg_flood_finish(int64_t n, const double* s, const double* W, double* route, pg_stream stream);

/* ---- the sea mask (bake/globe/erosion/maps.py:360-421 refresh_base, on
 * hydro/run.py:64-83 below_sea_components) -------------------------------
 * Replaces ErosionState.refresh_base on the globe (F = 6): base (DEVICE
 * float64, (6, NE, NE)) becomes 0 on the open ocean and on land and, on the
 * cells of a closed basin (a component of the cells whose float32 surface
 * height + sediment is < 0, 8-connected through owner, covering less than
 * the threshold and not the largest), the basin's lowest float64 surface;
 * halo cells copy their owner.  area_fx (DEVICE int64, (6, N, N)): each
 * interior cell's area as an exact fixed-point integer; threshold_fx: the
 * ocean area threshold (min_fraction x the globe's area) in the same units.
 * work: DEVICE scratch of pg_sea_work_bytes(N).  Host-synchronous (the
 * union-find's rounds are checked on the host).  info5 (HOST): closed cells,
 * closed basins, sea cells, deepest basin floor (cell units; 0 if none),
 * union-find rounds. */
int64_t pg_sea_work_bytes(int64_t N);
int pg_refresh_base(const double* height, const double* sediment, const int64_t* owner, const int64_t* area_fx,
                    double threshold_fx, int64_t N, int64_t H, double* base, void* work, int64_t work_bytes,
                    double* info5, pg_stream stream);

/* ---- fast mode (float32 droplets, order-independent fixed-point apply) -- */
/* Device pointers to the erosion state (bake/globe/erosion/maps.py:56-115
 * dtypes): the stage's fields in cell units. */
typedef struct pg_fast_fields {
    int64_t F, N, H;
    int64_t spherical, deposit_on_exit;
    double *height, *sediment, *discharge, *momentum, *pending;
    const double* base;
    const float *precip, *evap, *hardness, *metric, *metric_inv;
    const uint8_t* mask;
} pg_fast_fields;

/* The iteration's draws and knobs beyond pg_trace_params. */
typedef struct pg_fast_params {
    /* numpy PCG64 state after seeding with params.rng(stage, *key)
     * (bit_generator.state: state and inc, low / high 64-bit words) */
    uint64_t rng_state_lo, rng_state_hi, rng_inc_lo, rng_inc_hi;
    int64_t n_particles;   /* rain droplets: max(1, round(ppc * active cells)) */
    int64_t chunks;        /* frozen-terrain chunks per iteration (>= 1) */
    double iter_deposit;   /* net deposit a cell may take per iteration (cell units) */
    double iter_erode;     /* net erosion a cell may take per iteration (cell units) */
    double conc_chunk;     /* droplets per chunk the trace caps assume (the reference's chunk; 0 = `chunks`) */
    double fan_slope;      /* a submarine fan descends at least this much per cell (cell units per cell) */
    double offshore_writeoff;
    double ema;
    double min_volume_frac; /* death by evaporation below this x the spawn volume (prm->min_volume unused) */
    int64_t height_bits;   /* fixed point: heights in units of 2^-height_bits cells */
    int64_t track_bits;    /* fixed point: discharge / momentum in units of 2^-track_bits */
    int64_t grid_blocks;   /* persistent trace grid (0 = sized by occupancy) */
    int64_t time_trace;    /* record HIP events around every trace launch (see pg_fast_trace_time) */
    const uint64_t* rng_dev; /* optional DEVICE copy of the four rng words, read by the kernels instead of the
                              * values above: a captured graph then replays with each iteration's own draws */
    int64_t variant;       /* trace kernels' register budget: 0 = natural (census 4 waves per SIMD, traced applying pass 3), 1 / 2 = one / two waves more (same results) */
    const double* erode_cap; /* optional DEVICE per-cell erosion cap of the iteration (soillib's slope limit,
                              * pg_slope_erode_cap; bake/globe/erosion/maps.py:923-926), NULL = none */
    void* tape;            /* optional DEVICE move tape of pg_fast_tape_bytes() bytes: the census pass records
                            * every move and the applying pass replays it instead of tracing again (same
                            * results); NULL = trace twice.  pg_batch_iterate takes it from the first
                            * window's params, sized by pg_batch_tape_bytes */
    int64_t tape_bytes;
    const double* route;       /* optional DEVICE routing surface (F, NE, NE) (ErosionState.route, maps.py:351-358):
                                * droplets steer on it where it stands above the terrain and cross lakes;
                                * NULL = steer on the terrain */
    const uint8_t* lake_flag;  /* optional DEVICE lake flags (F, NE, NE) with the route (maps.py:629-672) */
    /* the discharge scales are pg_trace_params' disc_saturation and mom_scale, or, with
     * disc_saturation_cells / momentum_saturation_cells > 0, in cells of upstream rain, computed on the
     * device from the iteration's spawn (maps.py:911-921; inflow_volume = ErosionState.inflow_volume) */
    double disc_saturation_cells, momentum_saturation_cells, inflow_volume;
    /* The erosion brush (no counterpart in the reference, whose droplets change one cell per step,
     * particle.py:660-667): radius r >= 1 spreads each chunk's per-cell net height change over the
     * offsets |o| < r, weighted max(0, r - |o|), before the per-cell resolve (pg_brush_spread's rule);
     * radius 1 is the identity, 0 = off (the default).  brush_owner: DEVICE (F, NE, NE) int64 owner table
     * of the globe (bake/globe/cubesphere.py:665-686) through which halo shares reach their interior
     * cells; NULL on a window.  brush_tile: the stencil's LDS tile, 0 = 64 x 4, 1 = 32 x 8, 2 = 16 x 16
     * (same results).  pg_batch_iterate: every window's radius and tile are the first's, no owner. */
    int64_t brush_radius;
    const int64_t* brush_owner;
    int64_t brush_tile;
} pg_fast_params;

/* Bytes of the move tape for n_particles rain droplets on F faces of NE x NE
 * extended cells (plus at most one stockpile droplet per cell) in `chunks`
 * chunks of at most max_steps moves, for the persistent trace grid of
 * (grid_blocks, variant) as in pg_fast_params: per wave of that grid, rows of
 * 64 lane records of 32 bytes -- twice the wave's share of the largest
 * chunk's droplet-iterations plus one droplet's worth -- and a row count; and
 * a stop cell per droplet of the largest chunk. */
int64_t pg_fast_tape_bytes(int64_t n_particles, int64_t F, int64_t NE, int64_t chunks, int64_t max_steps,
                           int64_t grid_blocks, int64_t variant);
/* The same for a batch (pg_batch_iterate): every window's share of a chunk
 * (fps[w].n_particles and its extended cells, fps[0].chunks), max_steps the
 * windows' largest lifetime (which pg_batch_iterate reads from
 * prm->max_steps when a tape is given). */
int64_t pg_batch_tape_bytes(const pg_fast_fields* wins, const pg_fast_params* fps, int64_t n_win, int64_t max_steps,
                            int64_t grid_blocks, int64_t variant);

/* Statistics of one fast iteration (DEVICE int64[PG_FAST_NSTATS]). */
enum {
    PG_FS_PARTICLES = 0, PG_FS_RAIN, PG_FS_ENTRIES, PG_FS_MOVES, PG_FS_DEATHS, /* 6 causes: 4..9 */
    PG_FS_CLAMPED = 10, PG_FS_TO_PENDING_FX, PG_FS_LOST_OFFSHORE_FX, PG_FS_EXIT_FX, PG_FS_VANISHED_FX,
    PG_FS_RELEASED_FX, PG_FS_VOLUME0_BITS, PG_FS_ELIGIBLE, PG_FS_CDF_SERIAL, PG_FS_ERODED_FX, PG_FS_DEPOSITED_FX,
    PG_FS_DEFICIT_FX,
    PG_FS_LATERAL_FX = 22, /* the bank cuts' erosion taken (fixed point, erosion.lateral_rate) */
    PG_FS_TAPE_SHORT = 23, /* droplets a chunk's census left untraced, its tape rows full (always 0) */
    PG_FAST_NSTATS = 24
};

/* Bytes of DEVICE workspace the fast mode needs for this grid and rain
 * droplet count (fixed-point accumulators, packed records, spawn buffers,
 * sort/scan scratch).  Queries the current device's sort configuration. */
int64_t pg_fast_workspace_bytes(int64_t F, int64_t NE, int64_t n_particles);

/* Initialise a workspace for a state: static per-cell geometry records and
 * the fixed-point pending stockpile from fields->pending. */
int pg_fast_setup(const pg_fast_fields* fields, const pg_fast_params* fp, void* work, int64_t work_bytes,
                  pg_stream stream);

/* One iteration's particle pass (bake/globe/erosion/maps.py:839-1047) in
 * fast mode -- per chunk two trace passes over identical trajectories: an
 * erosion census, a per-cell erosion scale (iteration budget, base level,
 * lowest surface stepped to), then the applying pass where each droplet's
 * shortfall cancels its later deposits (plain profile: no routing surface, lakes, lateral cuts or
 * roots; prm->conc_scale and prm->min_volume are derived on the device), entirely on the device: pack, spawn (the params.rng stream:
 * sorted uniforms, inverse CDF, offsets), pending re-injection, `chunks`
 * rounds of persistent float32 trace + per-cell resolve against the live
 * terrain, then fold, EMA and the pending mirror.  Statistics go to the
 * DEVICE array stats (int64[PG_FAST_NSTATS]), zeroed first. */
int pg_fast_iterate(const pg_fast_fields* fields, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
                    int64_t work_bytes, int64_t* stats, pg_stream stream);

/* The two halves of pg_fast_iterate: spawn (reset statistics, pack, rain
 * and stockpile droplets) and the particle pass (chunks, finish). */
int pg_fast_spawn(const pg_fast_fields* fields, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
                  int64_t work_bytes, int64_t* stats, pg_stream stream);
int pg_fast_particles(const pg_fast_fields* fields, const pg_trace_params* prm, const pg_fast_params* fp, void* work,
                      int64_t work_bytes, int64_t* stats, pg_stream stream);

/* Byte offsets inside a fast workspace, for reading it: offsets[0..15] =
 * stencil records (float4), geometry (2 x float4), {sediment, base}
 * (float2), chunk / iteration / discharge / momentum / stockpile fixed
 * point (int64), droplet cell (int32), x / y fraction (float32), load
 * (int64), direction a / b (float32), scalars, total size. */
int pg_fast_layout(int64_t F, int64_t NE, int64_t n_particles, int64_t* offsets, int n);

/* Waits for the trace launches recorded since the last call (with
 * time_trace set) and returns their summed duration (ms, HIP events on the
 * launch stream) and count.  Library-wide state: not thread-safe. */
int pg_fast_trace_time(double* ms_total, int64_t* launches);

/* As pg_fast_trace_time, split by pass: ms[0] / launches[0] the census
 * launches (trace pass 0), ms[1] / launches[1] the applying pass (the
 * replay, or the traced applying pass without a tape).  Same reset. */
int pg_fast_trace_time_passes(double* ms, int64_t* launches);

/* The brush launches timed the same way (pg_fast_params.time_trace): total
 * ms and count since the last call; every timed launch read back so far is
 * kept per kind, so this and pg_fast_trace_time_passes may be called in
 * either order. */
int pg_fast_brush_time(double* ms, int64_t* launches);

/* ---- the erosion brush (no counterpart in the reference: its droplets
 * change one cell per step, bake/globe/erosion/particle.py:660-667) ------ */
/* The integer weights of a brush radius (0..4): (2A+1)^2 shares in units of
 * 2^-16, row-major over the offsets (dy, dx) in [-A, A]^2 with A = radius - 1
 * (A = 0 for radius 0 and 1), the centre's 0.  Returns A, or -1 when n is too
 * small or the radius out of range. */
int pg_brush_weights(int64_t radius, int32_t* w, int64_t n);
/* dst = src (int64 fixed-point amounts per extended cell, DEVICE, (F, NE,
 * NE)) spread over the brush footprint: every non-zero src[s] gives
 * sign(d) floor(|d| W_o / 2^16) to each valid target s + o (o != 0) in its
 * face's extended array and keeps the rest; a target is valid when it is an
 * active cell (mask 1) or a halo cell whose owner (owner table, DEVICE int64
 * (F, NE, NE); NULL = identity, a window) is active, and a halo cell's
 * shares are then added to its owner.  sum(dst) = sum(src) exactly.
 * valid: DEVICE scratch of F * NE * NE bytes (the target flags).  tile: 0,
 * 1, 2 as pg_fast_params.brush_tile. */
int pg_brush_spread(const int64_t* src, int64_t* dst, const uint8_t* mask, const int64_t* owner, int64_t F,
                    int64_t NE, int64_t radius, int64_t tile, uint8_t* valid, pg_stream stream);

/* ---- batched windows (no counterpart in the reference, which erodes its
 * zoom tiles one process each: bake/globe/zoom/bake.py:777-923) ----------
 * n_win single-face windows (F = 1, not spherical, same H and
 * deposit_on_exit) whose fields lie back to back in window order, one arena
 * per field.  fps[w]: window w's rng state, n_particles and conc_chunk; the
 * other knobs are fps[0]'s (chunks must agree).  prm: the shared trace
 * parameters (prm->max_steps ignored: max_steps[w] per window). */
int64_t pg_batch_workspace_bytes(int64_t n_win, const int64_t* NE, const int64_t* n_particles);
int pg_batch_setup(const pg_fast_fields* wins, int64_t n_win, const pg_fast_params* fps, const int64_t* max_steps,
                   void* work, int64_t work_bytes, pg_stream stream);
/* One fast-mode iteration of every window: per-window pack and spawn, then
 * per chunk both trace passes as one persistent launch over all windows'
 * droplets and the per-cell passes over the arena, then fold / EMA.
 * Bit-identical to pg_fast_iterate on each window alone.  stats: DEVICE
 * int64[PG_FAST_NSTATS] of the batch; win_stats: DEVICE
 * int64[n_win][PG_FAST_NSTATS] of the spawns. */
int pg_batch_iterate(const pg_fast_fields* wins, int64_t n_win, const pg_trace_params* prm, const pg_fast_params* fps,
                     void* work, int64_t work_bytes, int64_t* stats, int64_t* win_stats, pg_stream stream);

/* ---- differentiable mode (no counterpart in the reference) ---------------
 * The iteration as a function of the surface, a per-cell erodibility factor
 * and six scalars with the droplet trajectories frozen, and its adjoint.
 * Exchange arithmetic of pg_trace_particles' plain profile (particle.py:
 * 562-660) without apply-time caps, one frozen-terrain chunk per iteration. */
typedef struct pg_grad_args {
    int32_t real;               /* 0: float32 exchange arithmetic, 1: float64 */
    int32_t feedback;           /* reverse: 1 = propagate the surface adjoint through the exchange, 0 = hold the
                                 * surface the droplets read fixed (its adjoint passes the pass unchanged) */
    const void* surf;           /* exchange surface (height + sediment), F*NE*NE float or double */
    const float* erod;          /* per-cell erodibility factor multiplying c_eq, or NULL (1) */
    double theta[6];            /* erodibility, deposition_rate, evap_rate, k_disc, ocean_rate, dep_floor (cell units) */
    void* tape;                 /* NULL or pg_grad_tape_bytes() bytes: every move's exchange inputs */
    int64_t tape_bytes;
    const int32_t* fixed_moves; /* NULL, or per droplet (moves | cause << 24) of the run being replayed */
    int32_t* moves_out;         /* NULL, or per droplet (moves | cause << 24) of this run */
    void* rev_work;             /* reverse: DEVICE scratch of pg_grad_reverse_work_bytes() bytes (the
                                 * fixed-point accumulators that make the sums order-free) */
    int64_t rev_work_bytes;
    int64_t d_lo, d_hi;         /* forward: trace only the spawn's droplets [d_lo, d_hi), droplet d_lo + i on
                                 * tape slot i, and record the tape only: the target is not written
                                 * (d_hi <= d_lo: all droplets, applied).  The droplets of an iteration do not
                                 * interact, so an iteration's tape may be recorded and reversed in pieces */
    int32_t rev_flags;          /* reverse: bit 0 keeps the accumulators of the previous call, bit 1 leaves
                                 * them unflushed; pieces reversed with 2, 3, ..., 3, 1 give the whole
                                 * tape's gradients bit for bit (0: one call per tape) */
    int32_t reuse_spawn;        /* forward: 1 = the workspace still holds this iteration's packed record
                                 * and spawn (the previous call was the same iteration from the same
                                 * steering state): skip them */
} pg_grad_args;

/* Bytes of the reverse sweep's scratch for F faces of N x N cells, halo H:
 * a signed 128-bit fixed-point accumulator (unit 2^-64) per extended cell for
 * the surface and the erodibility adjoints, and eight for the scalars. */
int64_t pg_grad_reverse_work_bytes(int64_t F, int64_t N, int64_t H);

/* Bytes of a tape holding `capacity` droplets of up to max_steps moves. */
int pg_grad_tape_bytes(int64_t capacity, int64_t max_steps, int32_t real, int64_t* bytes);

/* One iteration's particle pass in the differentiable mode: pack and spawn
 * from `steer` (its packed surface, discharge and base steer the droplets;
 * fp->chunks must be 1), exchange against args->surf with args->theta and
 * args->erod, fold + EMA into `target`.  With args->fixed_moves every
 * droplet stops after the given moves with the given cause (the
 * trajectory of the run that wrote them). */
int pg_grad_iterate(const pg_fast_fields* steer, const pg_fast_fields* target, const pg_trace_params* prm,
                    const pg_fast_params* fp, const pg_grad_args* args, void* work, int64_t work_bytes,
                    int64_t* stats, pg_stream stream);

/* Reverse sweep of a taped iteration: with g = dL/d(surface after the
 * particle pass) (F*NE*NE double), adds the pass's share of dL/d(surface
 * before it) into lam, dL/d(erodibility factor) into grad_erod (F*NE*NE
 * double) and dL/d(theta) into grad_theta (double[6]).  Every contribution
 * is accumulated in 128-bit fixed point (args->rev_work; each truncated
 * toward zero at 2^-64, |contribution| < 2^62) and added to the outputs
 * once, so the result does not depend on the order of the lanes: bit for
 * bit reproducible. */
int pg_grad_reverse(int64_t F, int64_t N, int64_t H, const pg_trace_params* prm, const pg_grad_args* args,
                    const double* g, double* lam, double* grad_erod, double* grad_theta, pg_stream stream);

/* Fill out[k] = the (start + k)-th double of a PCG64 stream (numpy's
 * Generator.random() draws), k < n: the fast mode's spawn source. */
int pg_pcg64_random(uint64_t state_lo, uint64_t state_hi, uint64_t inc_lo, uint64_t inc_hi, int64_t start, int64_t n,
                    double* out, pg_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* PG_ERODE_H */
