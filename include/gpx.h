/*
 * gpx: C ABI of the MI355X FDTD solve path (libgprmax_hip.so).
 *
 * A gpx_ctx owns every device buffer of one model: the six field arrays, the
 * material-index array ID, the coefficient tables, PML auxiliary arrays,
 * dispersive auxiliary arrays, sources, receivers and snapshots.  Host arrays
 * passed in are borrowed for the duration of the call only and copied
 * synchronously.  All layouts are the reference's:
 *
 *   fields  float32 [nx+1][ny+1][nz+1], z fastest          (gprMax/grid.py:175-181)
 *   ID      uint32  [6][nx+1][ny+1][nz+1], Ex,Ey,Ez,Hx,Hy,Hz (gprMax/grid.py:167)
 *   coeffs  float32 [nmat][5] = (CA,CBx,CBy,CBz,srce) / (DA,DBx,DBy,DBz,srcm)
 *                                                           (gprMax/materials.py:200-201)
 *   disp    complex64 as float32 pairs [nmat][3*P][2] = (e0*eqt2, eqt, zt) per pole
 *                                                           (gprMax/materials.py:204-208)
 *   PML R   float32 [4][order][thickness] = RA, RB, RE, RF  (gprMax/pml.py:230-274)
 *   PML Phi float32 [order][A][B][C], the padded shapes of gprMax/pml.py:202-219
 *   rx buf  float32 [6][iterations][nrx]                    (gprMax/receivers.py:62)
 *
 * Every function returns 0 on success, a positive hipError_t for a HIP
 * failure or a negative GPX_E* code for a bad argument or state; the message
 * is available from gpx_last_error(ctx) (or gpx_last_error(NULL) for
 * gpx_create failures).  Nothing throws across the ABI.  A ctx is used from
 * one host thread at a time; all its work goes to one HIP stream.
 */
#ifndef GPX_H
#define GPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPX_EARG   (-1) /* invalid argument */
#define GPX_ESTATE (-2) /* call not valid in the current state */
#define GPX_ELIMIT (-3) /* exceeds a compiled-in limit */

typedef struct gpx_ctx gpx_ctx;

typedef struct gpx_desc {
    int nx, ny, nz;       /* grid size in cells (G.nx, G.ny, G.nz) */
    int iterations;       /* G.iterations */
    int nmat;             /* rows of updatecoeffsE/H */
    int maxpoles;         /* Material.maxpoles; 0 when no dispersive material */
    int pml_formulation;  /* 0 = HORIPML, 1 = MRIPML (G.pmlformulation) */
    int device;           /* HIP device ordinal */
} gpx_desc;

/* Sub-steps of one iteration, for gpx_part (gprMax/model_build_run.py:424-470). */
enum gpx_part_id {
    GPX_PART_STORE_RX = 0,   /* store_outputs, before the H update            */
    GPX_PART_SNAPSHOT = 1,   /* store snapshot `arg`                          */
    GPX_PART_H_MAIN = 2,     /* update_magnetic                               */
    GPX_PART_H_PML = 3,      /* PML.update_magnetic of slab `arg`             */
    GPX_PART_H_SRC = 4,      /* magnetic dipoles                              */
    GPX_PART_E_MAIN = 5,     /* update_electric, or dispersive part A         */
    GPX_PART_E_PML = 6,      /* PML.update_electric of slab `arg`             */
    GPX_PART_E_SRC = 7,      /* voltage sources then Hertzian dipoles         */
    GPX_PART_E_DISP_B = 8    /* dispersive part B                             */
};

/* Source kinds for gpx_set_sources. */
enum gpx_source_kind { GPX_SRC_HERTZIAN = 0, GPX_SRC_MAGNETIC = 1, GPX_SRC_VOLTAGE = 2 };

/* Context lifetime.  Replaces the device set-up of solve_gpu
 * (gprMax/model_build_run.py:477-583).  A grid of one cell along one axis is
 * a 2-D model (TMx, TMy or TMz, gprMax/input_cmds_singleuse.py:145-159);
 * more than one such axis is GPX_EARG. */
int gpx_create(const gpx_desc *desc, gpx_ctx **out);
int gpx_destroy(gpx_ctx *ctx);
const char *gpx_last_error(const gpx_ctx *ctx);

/* Model data.  Replace the host->device copies of grid.py:251-272,
 * model_build_run.py:512-519 and 540-549.  In a dispersive model the
 * auxiliary values (T, grid.py:187-190) are kept only for edges whose
 * material has poles; once both the indices and the dispersive tables are
 * on the device those edges are listed and their values set to zero, and a
 * later upload of either lists them again from zero. */
int gpx_upload_id(gpx_ctx *ctx, const uint32_t *ID);
int gpx_upload_coeffs(gpx_ctx *ctx, const float *coeffsE, const float *coeffsH);
int gpx_upload_dispersive(gpx_ctx *ctx, const float *disp_complex);

/* PML slabs, added in the order they are applied (x0, y0, z0, xmax, ymax,
 * zmax; gprMax/pml.py:307-335).  direction: 0 xminus, 1 yminus, 2 zminus,
 * 3 xplus, 4 yplus, 5 zplus.  Returns the slab index through *slab. */
int gpx_add_pml(gpx_ctx *ctx, int direction, int xs, int xf, int ys, int yf, int zs, int zf,
                int order, float d, const float *eR, const float *hR, int *slab);

/* Sources of one kind (gprMax/sources.py:235-283).  info int32 [n][6] =
 * (i, j, k, pol, first_it, last_it); pol + 4 marks a hard voltage source.
 * factors float32 [n][2]; waves float32 [n][iterations]. */
int gpx_set_sources(gpx_ctx *ctx, int kind, int n, const int32_t *info, const float *factors,
                    const float *waves);

/* A transmission line (gprMax/sources.py:286-451) attached to the E edge
 * (i, j, k) of polarisation pol; the reference runs these on the CPU only.
 * info int32 [9] = (i, j, k, pol, first_it, last_it, nl, srcpos, antpos);
 * coef float64 [4] = (R c dt/dl, (1/R) c dt/dl, c dt/dl, ABC factor
 * (c dt - dl)/(c dt + dl)); sizes float32 [3] = (cell size along pol, and the
 * two sizes of the current loop: dy, dz for x; dx, dz for y; dx, dy for z);
 * whole, half float32 [iterations] the waveform; voltage, current float32
 * [nl] and abc float32 [2] the line state after the incident run
 * (calculate_incident_V_I).  After the H pass the line records its antenna
 * voltage and current, then updates its current and takes the current of
 * the H loop around the edge; after the E pass it updates its voltage and
 * sets the edge's E.  Models with transmission lines step in the two-pass
 * form; dispersive models, a magnetic dipole on an H component the line
 * reads, or a Hertzian dipole on its edge are refused at the first step
 * (GPX_EARG).  Returns the line's index through *index. */
int gpx_add_transmission_line(gpx_ctx *ctx, const int32_t *info, const double *coef, const float *sizes,
                              const float *whole, const float *half, const float *voltage,
                              const float *current, const float *abc, int *index);

/* The antenna-end voltage and current of every iteration (Vtotal, Itotal,
 * float32 [iterations]) and the line's voltage and current now (float32
 * [nl]); any pointer may be NULL.  Waits for the ctx stream. */
int gpx_download_transmission_line(gpx_ctx *ctx, int index, float *vtotal, float *itotal, float *voltage,
                                   float *current);

/* Receivers (gprMax/receivers.py:45-68): coords int32 [n][3]. */
int gpx_set_receivers(gpx_ctx *ctx, int n, const int32_t *coords);

/* Snapshots (gprMax/snapshots.py:170-201): start cell, step, output size,
 * and the iteration+1 at which it is taken.  Returns its index. */
int gpx_add_snapshot(gpx_ctx *ctx, int xs, int ys, int zs, int dx, int dy, int dz,
                     int snx, int sny, int snz, int time, int *index);

/* The time loop: iterations [it0, it0 + n), every sub-step of an iteration
 * inside the kernels of the step form set by gpx_set_step (replaces
 * model_build_run.py:590-696).  Enqueued on the ctx stream; returns without
 * waiting, except that under the default form the first run of at least 26
 * iterations waits once, after its 14th iteration, to read its timing. */
int gpx_run(gpx_ctx *ctx, int it0, int n);

/* As gpx_run, and reports the elapsed device time measured with HIP events
 * on the ctx stream, plus the mean duration of the H and E passes when
 * per_kernel is non-zero (events between every launch).  Waits. */
int gpx_run_timed(gpx_ctx *ctx, int it0, int n, int per_kernel, float *ms_total,
                  float *ms_h, float *ms_e);

/* As gpx_run, with HIP events around every kernel; ms4 receives the mean
 * duration in ms of each kernel of an iteration: the fused step kernel in
 * ms4[0] (ms4[1..3] = 0); the split step's interior kernel, H boundary and E
 * boundary passes in ms4[0], ms4[1], ms4[3] (ms4[2] = 0); or with the
 * two-pass form H interior, H boundary, E interior, E boundary (interior =
 * cells away from the faces and PML slabs, boundary = the rest).  The form 4
 * choice is not timed here: a profiled run uses the split step until a
 * gpx_run has made the choice.  Waits once, at the end. */
int gpx_run_profiled(gpx_ctx *ctx, int it0, int n, float *ms4);

/* Cells of the interior box (the rest are boundary cells), and the box
 * itself as [x0, x1, y0, y1, z0, z1), cell indices of the field arrays. */
long long gpx_interior_cells(const gpx_ctx *ctx);
int gpx_interior_box(const gpx_ctx *ctx, int *box6);

/* One unfused sub-step (enum gpx_part_id) at iteration it. */
int gpx_part(gpx_ctx *ctx, int part, int arg, int it);

/* Launch shape: planes per chunk along x (0 = auto, both step forms) and
 * threads per block of the H and E passes (64, 128, 256, 512 or 1024; 0 = auto). */
int gpx_set_launch(gpx_ctx *ctx, int xchunk, int block);

/* Step form of gpx_run.  fused = 1: one kernel per iteration that does the
 * H and the E update with the state double-buffered on the device; fused = 0:
 * an H pass and an E pass per iteration; fused = 2: the passes with their
 * interior and boundary kernels running side by side on two streams, one
 * block per item each (gpx_run_profiled always runs them one after the
 * other); fused = 3: the
 * split step, one kernel doing H and E for the interior and the H and E
 * boundary passes around it, double-buffered; fused = 5: the split step with
 * the H boundary pass running beside the interior kernel on a second stream
 * (gpx_run_profiled runs them one after the other); fused = 6: the passes,
 * each as one launch over every cell, the interior and boundary items of
 * each x chunk next to each other; fused = 7: the split step with the H
 * boundary pass's items inside the interior kernel's launch, likewise
 * interleaved; fused = 4 (default): the first gpx_run of at least 26
 * iterations times forms 7, 6, 2, 5, the split step and the passes (forms 6,
 * 2 and the passes where the split step does not apply) on four iterations
 * each and keeps the fastest (gpx_step_choice).  tile_cells = cells per fused block (256, 512 or 1024;
 * 0 = auto).  All forms give identical results; the
 * fused form needs a tile of at least twice nz+1 cells within the LDS budget,
 * the split step a model without dispersive poles, and the passes run when
 * the chosen form does not apply (form 5 as the split step). */
int gpx_set_step(gpx_ctx *ctx, int fused, int tile_cells);

/* Cells per block of the fused step that gpx_run will use, or 0 when it will
 * run another form. */
int gpx_step_kind(const gpx_ctx *ctx);

/* The form gpx_run will use: 0 passes, 1 fused step, 2 passes side by side,
 * 3 split step, 5 split step with the H boundary pass beside the interior
 * kernel, 6 merged passes, 7 split step with the H boundary pass in the
 * interior kernel's launch.  gpx_run_profiled reports the split step's
 * interior kernel in ms4[0] and its boundary passes in ms4[1] and ms4[3]
 * (form 7: the merged launch in ms4[0], ms4[1] = 0), and the merged passes'
 * H and E launches in ms4[0] and ms4[2]. */
int gpx_step_form(const gpx_ctx *ctx);

/* Form 4's choice: -1 not timed yet, 0 the passes, 1 the split step, 2 form 5,
 * 3 form 2, 4 form 6, 5 form 7. */
int gpx_step_choice(const gpx_ctx *ctx);

int gpx_synchronize(gpx_ctx *ctx);

/* State transfer (synchronous). */
int gpx_download_rx(gpx_ctx *ctx, float *out);
int gpx_download_fields(gpx_ctx *ctx, float *out6);
int gpx_upload_fields(gpx_ctx *ctx, const float *in6);
int gpx_download_pml_phi(gpx_ctx *ctx, int slab, int which, float *out);
int gpx_upload_pml_phi(gpx_ctx *ctx, int slab, int which, const float *in);
/* T in the reference's layout, complex64 [3][P][nx+1][ny+1][nz+1]; edges
 * without poles read back as zero and their uploaded values are ignored. */
int gpx_download_dispersive_T(gpx_ctx *ctx, float *out_complex);
int gpx_upload_dispersive_T(gpx_ctx *ctx, const float *in_complex);
int gpx_download_snapshot(gpx_ctx *ctx, int index, float *out6);

/* Device query for the drop-in GPU object (gprMax/utilities.py:341-366). */
int gpx_device_info(int device, char *name, int namelen, size_t *totalmem, size_t *freemem,
                    int *cus);
/* Bytes of device memory held by ctx. */
size_t gpx_memory_bytes(const gpx_ctx *ctx);

/* ---- Batched B-scan: many models advanced by shared launches.
 *
 * Replaces the reference's sequential trace loop, in which every B-scan trace
 * re-runs run_model and so a whole solve_gpu (gprMax/gprMax.py:222-232 with
 * gprMax/model_build_run.py:84-406); here the traces stay separate gpx_ctx
 * models, each built and uploaded as usual.  Each iteration of the batch is
 * either four launches (H interior, H boundary, E interior, E boundary), each
 * over the work of every model (tile 0, the default), or one launch of the
 * fused step over every model (tile > 0).  Requirements: one device, the same
 * iteration count, no snapshots, and the same kernel (index width, poles and
 * their kind, PML formulation) for every model; the fused step also needs one
 * PML order for all slabs of all models.  The contexts must outlive the
 * batch; results are read from each context (gpx_download_rx, ...). */
typedef struct gpx_batch gpx_batch;

int gpx_batch_create(gpx_ctx *const *ctxs, int n, gpx_batch **out);
int gpx_batch_destroy(gpx_batch *b);
/* Message of the last failure (NULL: of the last gpx_batch_create). */
const char *gpx_batch_last_error(const gpx_batch *b);
/* tile_cells 0: the H and E passes; 256, 512 or 1024: the fused step with
 * that many cells per block and `xchunk` planes per x chunk (0 = chosen for
 * the batch). */
int gpx_batch_set_shape(gpx_batch *b, int tile_cells, int xchunk);
/* Autotuner: times `steps` iterations of the passes and of the fused step for
 * each usable tile size and three chunk lengths around the planned one, keeps
 * the fastest (*tile_cells 0 for the passes), reports it and its time per
 * iteration, then resets every model to the zero state (fields, Phi, T,
 * receiver buffer).  Call before stepping. */
int gpx_batch_tune(gpx_batch *b, int steps, int *tile_cells, int *xchunk, float *ms_per_step);
/* The passes advance the models in groups of consecutive models whose state
 * (fields, indices, Phi, T) adds up to at most `bytes` (0: 192 MiB, to stay in
 * the 256 MB last-level cache; a larger model forms its own group); each group
 * runs all its iterations before the next starts. */
int gpx_batch_set_group_bytes(gpx_batch *b, double bytes);
/* Iterations [it0, it0 + n) of every model; graph != 0 captures the n
 * launches in one HIP graph and launches that.  Waits; *ms (optional) is the
 * device time measured with HIP events around the launches. */
int gpx_batch_run(gpx_batch *b, int it0, int n, int graph, float *ms);

/* ---- Differentiable solve: forward with checkpoints and its adjoint.
 *
 * No counterpart in the reference (gprMax has no gradients; its optimiser is
 * gradient-free, gprMax/optimisation_taguchi.py).  The forward scheme is the
 * reference's solver loop (gprMax/model_build_run.py:424-470) with every cell
 * its own material: per-cell CA, CB = CBx = CBy = CBz (cubic cells) and
 * IEA = srce = 1/EA (gprMax/materials.py:178-201), mu_r = 1 and no magnetic
 * loss, HORIPML with one CFS term, one z-directed Hertzian dipole and Ez
 * receivers.  gpx_inv_backward returns dL/dCA, dL/dCB and dL/dIEA per cell
 * for L with dL/d(trace) = resid; the PML correction's use of IEA is not
 * differentiated, so gradients of cells inside a PML slab are incomplete. */
typedef struct gpx_inv gpx_inv;

typedef struct gpx_inv_desc {
    int nx, ny, nz;       /* cells */
    int iterations;
    int K;                /* checkpoint interval (iterations) */
    int device;
    float d;              /* cell size (dx = dy = dz) */
    float dt;             /* time step */
} gpx_inv_desc;

int gpx_inv_create(const gpx_inv_desc *d, gpx_inv **out);
int gpx_inv_destroy(gpx_inv *v);
const char *gpx_inv_last_error(const gpx_inv *v);
/* One HORIPML slab (order 1), as gpx_add_pml; eR, hR = RA, RB, RE, RF [t]. */
int gpx_inv_add_pml(gpx_inv *v, int direction, int xs, int xf, int ys, int yf, int zs, int zf,
                    const float *eR, const float *hR);
/* Per-cell coefficients, float32 [nx][ny][nz] each. */
int gpx_inv_set_model(gpx_inv *v, const float *CA, const float *CB, const float *IEA);
/* Debye poles per cell with real auxiliary values, the reference's dispersive
 * parts A and B (gprMax/fields_updates_ext.pyx:113-235) on every updated E
 * edge: P poles (0 removes them, at most 4), and per pole float32
 * [P][nx][ny][nz] the coefficients ee = e0 Re(eqt2), eqt and zt of
 * gprMax/materials.py:92-143 (zeros in cells without poles).  The poles are
 * fixed: the gradient stays with respect to CA, CB and IEA, whose EA the
 * caller forms with the -(e0/dt) sum Re(zt2) term of materials.py:127-128. */
int gpx_inv_set_dispersive(gpx_inv *v, int P, const float *ee, const float *eqt, const float *zt);
/* All iterations from the zero state with a z-directed Hertzian dipole on
 * the Ez edge (si, sj, sk), Ez -= IEA * waves[n] * sfac; receivers rx int32
 * [nrx][3]; traces float32 [nrx][iterations] (Ez at the start of each
 * iteration).  Keeps a checkpoint of the state every K iterations. */
int gpx_inv_forward(gpx_inv *v, int si, int sj, int sk, float sfac, const float *waves, int nrx,
                    const int *rx, float *traces);
/* Adjoint of the last forward: resid float32 [nrx][iterations] = dL/d(trace);
 * outputs float32 [nx][ny][nz].  Recomputes each checkpoint segment. */
int gpx_inv_backward(gpx_inv *v, const float *resid, float *gCA, float *gCB, float *gIEA);
/* Device memory the forward may spend keeping the stored states (E^n, the
 * curl of H^{n+1}, and phi^n with poles: 6 or 9 x (nx+1)(ny+1)(nz+1) floats per
 * iteration) of its last checkpoint segments, so that the backward does not
 * recompute them (default 0: every segment is recomputed).  As many whole
 * segments as fit are kept, from the last one down; the gradient is the same
 * bit for bit.  Takes effect at the next gpx_inv_forward.  No counterpart in
 * the reference. */
int gpx_inv_set_keep(gpx_inv *v, double bytes);
/* Device time of the last forward and of the last backward (recompute
 * included), ms. */
int gpx_inv_timings(const gpx_inv *v, float *ms_forward, float *ms_backward);

/* ---- Edge material indices on the device.
 *
 * Replaces the edge loops of build_electric_components and
 * build_magnetic_components (gprMax/yee_cell_build_ext.pyx:110-257; called at
 * gprMax/model_build_run.py:215-219) for grids whose cell materials are below
 * 65535.  solid uint32 [nx][ny][nz] (cell materials), rigidE / rigidH uint8
 * [nx][ny][nz] (cells whose edges the geometry fixed), ID uint32
 * [6][nx+1][ny+1][nz+1] (the indices the geometry wrote).  Per field (0: E,
 * then 1: H, the reference's order): gpx_edges_classify writes the edges whose
 * cells agree and returns the distinct sets of cell materials of the others
 * (materials sorted and packed 16 bits each, with the first edge of each set
 * as component * (nx+1)(ny+1)(nz+1) + point index); the caller finds or
 * creates each set's averaged material in first-edge order
 * (create_electric_average / create_magnetic_average,
 * gprMax/yee_cell_build_ext.pyx:31-107) and passes them, in the order the sets
 * came, to gpx_edges_resolve, which writes them to every edge of the set.
 * `capacity` bounds the distinct sets per field (GPX_ELIMIT beyond). */
typedef struct gpx_edges gpx_edges;

int gpx_edges_create(int device, int nx, int ny, int nz, const uint32_t *solid, const uint8_t *rigidE,
                     const uint8_t *rigidH, const uint32_t *ID, int capacity, gpx_edges **out);
int gpx_edges_destroy(gpx_edges *b);
const char *gpx_edges_last_error(const gpx_edges *b);
int gpx_edges_classify(gpx_edges *b, int field, unsigned long long *keys, unsigned long long *first, int max,
                       int *count);
int gpx_edges_resolve(gpx_edges *b, int field, const uint32_t *numids, int count);
int gpx_edges_download(gpx_edges *b, uint32_t *ID);

/* ---- Edge build that allocates nothing.
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the work of gpx_edges_classify and gpx_edges_resolve on
 * the caller's device arrays, with the hash tables in a fixed 3.1 MiB device
 * array of the library (one per GPU, kept until the process ends) instead of
 * allocated memory.  solid, rigid (rigidE for field 0, rigidH for field 1) and
 * ID are DEVICE pointers on `device`, laid out as for gpx_edges_create; keys,
 * first, known_keys, known_ids and count are host pointers.  One call writes
 * the field's uniform edges, writes the mixed edges whose set is among the
 * `nknown` (known_keys -> known_ids) pairs, and returns the other sets as
 * gpx_edges_classify does (at most `max` written, *count their number, table
 * order).  The usual sequence per field: call with nknown = 0, resolve the
 * returned sets in first-edge order, call again with them (then *count is
 * 0).  At most 32768 sets per call, given or returned (GPX_ELIMIT beyond).
 * Runs on the default stream and returns when the work is done; calls that
 * overlap on one GPU share the array and race. */
int gpx_edges_static(int device, int nx, int ny, int nz, const uint32_t *solid, const uint8_t *rigid,
                     uint32_t *ID, int field, int nknown, const unsigned long long *known_keys,
                     const uint32_t *known_ids, unsigned long long *keys, unsigned long long *first, int max,
                     int *count);

#ifdef __cplusplus
}
#endif

#endif /* GPX_H */
