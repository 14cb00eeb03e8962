/*
 * rca_das.h — C ABI of the MI355X (gfx950) RCA delay-and-sum I/Q beamformer.
 *
 * Plain C: device pointers, sizes and scalars; no torch types.  Every call is
 * asynchronous on the given HIP stream and allocates no device memory.
 * Calls on distinct streams are re-entrant.  Errors return a non-zero
 * rca_status and leave a message in rca_last_error() (thread-local).
 *
 * Layouts (all C-order, complex = interleaved float re/im):
 *   iq       complex64 [nFrame][nTx][nChannel][nSample]   (RCA_IQ_C64)
 *            or half2  [nFrame][nTx][nChannel][nSample] with a per-trace
 *            float scale [nFrame][nTx][nChannel]           (RCA_IQ_F16X2)
 *            == the reference's data [2*nSample, nChannel, nTx, nFrame]
 *            column-major buffer (bfVirtualSourceGeneric_RCA.m:576-578).
 *   out      complex64 [nFrame][1 | nTx][nPix]  (1 when RCA_FLAG_COMPOUND),
 *            or [nChannel][nFrame][nTx][nPix] with RCA_FLAG_NOSUM
 *            == the reference's dataBf [2*nPix, nTx|1, nFrame]
 *            (bfVirtualSourceGeneric_RCA.cu:230, .m:654-658).
 *   mask     float [1 | nTx][nPix] (sum of Hann weights, .cu:301-304).
 */
#ifndef RCA_DAS_H
#define RCA_DAS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  RCA_OK = 0,
  RCA_ERR_INVALID = 1,     /* bad argument (null pointer, size, flag) */
  RCA_ERR_UNSUPPORTED = 2, /* shape outside what the kernels handle */
  RCA_ERR_HIP = 3,         /* a HIP runtime call failed */
  RCA_ERR_BUDGET = 4       /* a request exceeds the serving memory budget */
} rca_status;

typedef enum {
  RCA_INTERP_QUADRATIC = 0,   /* reference default (.m:159-161; .cu:355-365) */
  RCA_INTERP_LANCZOS_REF = 1, /* reference "Lanczos 5-lobes" (.cu:331-350) */
  RCA_INTERP_LANCZOS3 = 2,    /* textbook Lanczos a=3, 6 taps */
  RCA_INTERP_LANCZOS4 = 3     /* textbook Lanczos a=4, 8 taps */
} rca_interp;

typedef enum { RCA_IQ_C64 = 0, RCA_IQ_F16X2 = 1 } rca_iq_dtype;

enum {
  RCA_FLAG_COMPOUND = 1,   /* sum over tilts in the kernel (.m:690-697) */
  RCA_FLAG_ACCUMULATE = 2, /* out += result (adds an RCA half, script_All.m:382) */
  RCA_FLAG_NOSUM = 4       /* one output per receive line, out [nChannel][nFrame][nTx][nPix]
                              (NoSum/bfVirtualSourceGeneric_RCA_Nosum.cu:228-229, 376-380) */
};

/*
 * Receive aberration laws: delays in seconds added to the receive time of
 * flight (and so to the sample position and the phase).
 *   CHANNEL        ab_law_rx [nChannel]          (Hyp1_save/bfVirtualSourceGeneric_RCA_ab_1.cu:306-311)
 *   TX_CHANNEL     ab_law_rx [nTx][nChannel]     (Hyp2/bfVirtualSourceGeneric_RCA_ab_2.cu:306-311)
 *   PIXEL_CHANNEL  ab_law_rx and ab_law_tx [nChannel][nPix], either may be
 *                  NULL (exemple_paul/bfVirtualSourceGenericTxRx.cu:268-275,
 *                  292-294: AbLawRx[idxPix + idxChannel*nPix] joins the
 *                  receive delay, AbLawTx[...] the total delay)
 */
typedef enum {
  RCA_ABLAW_NONE = 0,
  RCA_ABLAW_CHANNEL = 1,
  RCA_ABLAW_TX_CHANNEL = 2,
  RCA_ABLAW_PIXEL_CHANNEL = 3
} rca_ablaw;

/*
 * Beamforming geometry and parameters of one RCA half: the argument list of
 * the reference kernel (bfVirtualSourceGeneric_RCA.cu:103-133) with device
 * pointers, grouped so it can stay resident between calls.
 */
typedef struct {
  const float* grid_x;       /* [nPix] gridBfX (m) */
  const float* grid_y;       /* [nPix] gridBfY */
  const float* grid_z;       /* [nPix] gridBfZ */
  const float* ele_x;        /* [nChannel] centred posEleX */
  const float* ele_y;        /* [nChannel] centred posEleY */
  const float* ele_z;        /* [nChannel] centred posEleZ */
  const float* vs_x;         /* [nTx] virtualSourceX */
  const float* vs_y;         /* [nTx] virtualSourceY */
  const float* vs_z;         /* [nTx] virtualSourceZ */
  const float* min_dist;     /* [nTx] minDistVirtualSource */
  const float* ab_law_rx;    /* receive aberration delays (s) or NULL */
  int ab_law_mode;           /* rca_ablaw: [nChannel], [nTx][nChannel] or [nChannel][nPix] */
  int n_pix;
  int n_channel;
  int n_tx;
  int n_sample;              /* I/Q samples per trace, at most 2^27 */
  int receive_x;             /* flagReceiveX */
  int mask_transmit;         /* flagMaskTransmit */
  int interp;                /* rca_interp */
  float aperture_x, aperture_y, aperture_z;
  float sampling_frequency;
  float demodulation_frequency;
  float speed_of_sound;
  float start_time;
  float fnumber;
  /* Voxel layout hint, no counterpart in the reference: when the voxel list is
   * a grid stored depth-fastest (voxel p = iz + depth_run * column, MATLAB's
   * [Nz, Nx, Ny] order) pass Nz here (>= 64) and each workgroup takes a tile
   * of 64 depths x a few columns, which narrows the sample window it stages
   * per (tilt, line).  0 = no hint (workgroups take consecutive voxels).  The
   * results do not depend on it. */
  int depth_run;
  /* RCA_ABLAW_PIXEL_CHANNEL only: the transmit-side law [nChannel][nPix] (s),
   * or NULL (AbLawTx of bfVirtualSourceGenericTxRx.cu:122, 275, 294). */
  const float* ab_law_tx;
} rca_das_desc;

/*
 * Reference kernel ABI.  Same arguments, order and meaning as
 * bfVirtualSourceGeneric_RCA (MATLAB/lib/beamformer/RCA/bfVirtualSourceGeneric_RCA.cu:103-133,
 * launched by feval at bfVirtualSourceGeneric_RCA.m:670-686), all pointers on
 * the device, plus the stream.  Writes the per-tilt dataBf
 * [2*nPix*nTx*nFrame] and mask [nPix*nTx]; the caller sums over tilts when
 * compounding, as the reference wrapper does.  Pixels outside the transmit
 * cone are written as zero (the reference leaves them untouched).
 */
int bfVirtualSourceGeneric_RCA(float* dataBf, float* mask, const float* data,
                               const float* gridBfX, const float* gridBfY, const float* gridBfZ,
                               const float* posEleX, const float* posEleY, const float* posEleZ,
                               const float* virtualSourceX, const float* virtualSourceY,
                               const float* virtualSourceZ, const float* minDistVirtualSource,
                               int flagReceiveX, float apertureX, float apertureY, float apertureZ,
                               float samplingFrequency, float demodulationFrequency,
                               float speedOfSound, float startTime, float fnumber, int nPix,
                               int nChannel, int nTx, int nFrame, int nSample,
                               int flagInterpolationLanczos, int flagMaskTransmit,
                               hipStream_t stream);

/*
 * Forward beamforming of nFrame frames of one RCA half.
 * Replaces the feval + tilt-sum + gather sequence of
 * bfVirtualSourceGeneric_RCA.m:664-711.  trace_scale is required for
 * RCA_IQ_F16X2 and ignored for RCA_IQ_C64.  mask may be NULL.
 */
int rca_forward(const rca_das_desc* desc, const void* iq, int iq_dtype, const float* trace_scale,
                int n_frame, void* out, float* mask, int flags, hipStream_t stream);

/*
 * Adjoint (no reference counterpart; serves autograd through the forward):
 * grad_iq = A^H grad_out, complex64 [nFrame][nTx][nChannel][nSample].
 * grad_out is complex64 [nFrame][grad_tx][nPix] with grad_tx in {1, nTx}
 * (1 = gradient of the compounded output).  grad_iq is overwritten unless
 * RCA_FLAG_ACCUMULATE is set.
 */
int rca_adjoint(const rca_das_desc* desc, const void* grad_out, int grad_tx, int n_frame,
                void* grad_iq, int flags, hipStream_t stream);

/*
 * Deterministic adjoint (no reference counterpart; for training runs that
 * must reproduce bit for bit): the same A^H grad_out as rca_adjoint, with
 * the tiles' partial sums added as 64-bit fixed-point integers (2^-S quanta,
 * S set from n_pix and the largest component of grad_out so no sum can
 * overflow; about 1e-12 of that component) in `workspace`, then converted to
 * FP32 into grad_iq (overwritten, or added to with RCA_FLAG_ACCUMULATE).
 * Integer addition does not depend on the order in which tiles finish, so
 * repeated calls on the same inputs give identical grad_iq.  workspace:
 * device memory, 256-byte aligned, of rca_adjoint_det_workspace_bytes bytes
 * (256 + 16 per complex gradient sample); the call allocates nothing.  A
 * non-finite grad_out gives NaN throughout grad_iq.
 */
size_t rca_adjoint_det_workspace_bytes(const rca_das_desc* desc, int n_frame);
int rca_adjoint_det(const rca_das_desc* desc, const void* grad_out, int grad_tx, int n_frame,
                    void* grad_iq, int flags, void* workspace, size_t workspace_bytes,
                    hipStream_t stream);

/*
 * Batched forward for serving (no reference counterpart; the reference makes
 * one synchronous call per chunk, bfVirtualSourceGeneric_RCA.m:664-711).
 * A batch is a list of entries; each entry beamforms frames
 * [frame_begin, frame_end) of one request: one or two RCA halves on the same
 * voxel grid, compounded over tilts and summed over halves into `out`
 * (complex64 [n_frame][1][nPix]) and `mask` ([1][nPix], written by the batch
 * that contains frame 0).  rca_batch_encode fills a host workspace (header,
 * per-half arguments, work items); the caller copies it to a device buffer of
 * the same size; rca_batch_launch runs the batch as eight launches of
 * max_items workgroups, one work item per workgroup (max_items as given to
 * rca_batch_encode): for each item type - complex64 or FP16 I/Q, with or
 * without a depth hint (rca_das_desc.depth_run >= 64) - the first halves,
 * then the second halves.  The launch arguments never change
 * between batches, so one captured HIP graph replays every batch.  Per-voxel
 * arithmetic is the same as rca_forward's: results are bitwise identical
 * whatever else shares the batch.  Frames are processed in blocks of
 * rca_batch_frame_block(iq_dtype, depth_tiled): an entry's frame_begin must be
 * a multiple of it, and frame_end a multiple of it or n_frame.
 */
typedef struct {
  const rca_das_desc* half[2];
  int n_half;                  /* 1 or 2 */
  const void* iq[2];           /* per half, [n_frame][nTx][nChannel][nSample] */
  const float* trace_scale[2]; /* per half, FP16 only */
  int iq_dtype;                /* rca_iq_dtype */
  int n_frame;
  int frame_begin;
  int frame_end;
  void* out;
  float* mask;                 /* may be NULL */
} rca_batch_entry;

size_t rca_batch_workspace_bytes(int max_halves, int max_items);
int rca_batch_encode(const rca_batch_entry* entries, int n_entries, int max_halves, int max_items,
                     int n_sample_max, void* host_ws, size_t ws_bytes, int* n_items);
int rca_batch_launch(const void* dev_ws, int max_halves, int max_items, int interp,
                     int n_sample_max, hipStream_t stream);
/*
 * rca_batch_launch without a caller-allocated workspace (no reference
 * counterpart; for callers that cannot allocate device memory: code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Takes the work list as rca_batch_encode wrote it in host memory
 * (host_ws, ws_bytes >= rca_batch_workspace_bytes(max_halves, max_items)),
 * copies it on `stream` into a device array the library holds per GPU
 * (4 MiB: rca_batch_workspace_bytes(max_halves, max_items) must not exceed
 * it, else RCA_ERR_UNSUPPORTED) and runs the batch from there.  Allocates
 * nothing.  Calls that overlap on one GPU share that array and race; calls on
 * one stream are ordered.
 */
int rca_batch_launch_noalloc(const void* host_ws, size_t ws_bytes, int max_halves, int max_items,
                             int interp, int n_sample_max, hipStream_t stream);
/* Frames per block of a batch entry of this I/Q dtype, with (depth_tiled != 0)
 * or without a depth hint: 8 with one, else 1 for complex64 and 4 for FP16;
 * 0 for an unknown dtype. */
int rca_batch_frame_block(int iq_dtype, int depth_tiled);

/*
 * Launch-shape override for tuning sweeps (no reference counterpart):
 * voxels per lane v in {1,2,4} (forward and adjoint), frames per workgroup fb in {1,2,4,8} (forward
 * and adjoint), split: reserved (>= 0, ignored), LDS budget of a workgroup in
 * KiB (16..64, which sets how many tilts share one pass over the receive
 * lines); 0 restores the automatic choice.  Process-wide.
 */
int rca_set_tuning(int v, int fb, int split, int lds_kib);

/*
 * Resident plans (the boundary's rca_plan_create / rca_plan_destroy): the
 * host prep of bfVirtualSourceGeneric_RCA.m:491-558 in FP64 and one upload
 * of the geometry, so a C or MATLAB caller needs nothing but host arrays.
 * Arrays are host memory in MATLAB's column-major flatten; values are
 * rounded to FP32 as the wrapper's single() does (RCA.m:598-636).
 */
typedef struct {
  const double *grid_x, *grid_y, *grid_z;   /* [nPix] gridBfX/Y/Z (m) */
  const double *ele_x, *ele_y, *ele_z;      /* [nChannel] posEleX/Y/Z, as given (centred here) */
  const double *theta_x, *theta_y;          /* [nTx] tilt angles (rad), or both NULL ... */
  const double *vs_x, *vs_y, *vs_z;         /* ... and the virtual sources given instead */
  const double *ab_law_rx, *ab_law_tx;      /* aberration laws (s), layouts of rca_ablaw, or NULL */
  int n_pix, n_channel, n_tx, n_sample;
  int receive_x, mask_transmit, interp, ab_law_mode;
  int depth_run;                            /* as in rca_das_desc (0 = none) */
  int has_aperture;                         /* 0: apertures = span of the centred elements */
  double aperture_x, aperture_y, aperture_z;
  double sampling_frequency, demodulation_frequency, speed_of_sound, start_time, fnumber;
} rca_plan_params;

typedef struct rca_plan rca_plan;

/* Host-only prep (no GPU): centred elements [3][nChannel], virtual sources
 * [3][nTx] (tilts -> r = 10 m, RCA.m:527-541), minDist [nTx] with the
 * reference's X/Y swap (RCA.m:551-558) and apertures [3], all FP32. */
int rca_prepare_geometry(const rca_plan_params* p, float* ele, float* vs, float* min_dist,
                         float* aperture);

/* Allocate and upload a plan on the current device (the only allocation;
 * rca_forward / rca_adjoint on its descriptor allocate nothing). */
int rca_plan_create(const rca_plan_params* p, rca_plan** plan);
const rca_das_desc* rca_plan_desc(const rca_plan* plan);
int rca_plan_destroy(rca_plan* plan);

/*
 * Measurement probe (no reference counterpart): VALU issue ceilings the DAS
 * kernels are reported against.  n_blocks workgroups of 256 lanes, each lane
 * 8 independent chains x 8 x iters instructions of one class: mode 0
 * v_fma_f32 (VGPR multiplier), 1 v_fma_f32 (inline constants), 2 v_add_f32,
 * 3 v_mul_f32, 4 v_fma_f64, 5 v_max_f32, 6 v_add_u32, 7 v_fma_f32 with three
 * VGPR sources, 8 v_add_f32 and 9 v_mul_f32 with two, 10 v_fma_f32 with an
 * SGPR and two VGPRs, 11 v_fma_f32 with two VGPRs and a literal.  Writes one float per
 * lane to out (n_blocks * 256) and the lane-operations of the launch to
 * *lane_ops.  Asynchronous on stream.
 */
int rca_probe_valu(int mode, float* out, int n_blocks, int iters, hipStream_t stream,
                   long long* lane_ops);

/*
 * FP16-I/Q storage (no reference counterpart: the reference keeps I/Q in
 * FP32, bfVirtualSourceGeneric_RCA.m:476-484).  iq: complex64 traces
 * [n_trace][n_sample]; out: half2 [n_trace][n_sample]; scale: float
 * [n_trace].  Per trace, scale = max(|re|, |im|) over its samples (1 for an
 * all-zero trace) and each sample is stored as (re, im) / scale rounded to
 * the nearest half.  Asynchronous on stream.
 */
int rca_quantize_fp16(const void* iq, long long n_trace, int n_sample, void* out, float* scale,
                      hipStream_t stream);

/* Last error message of the calling thread ("" when none). */
const char* rca_last_error(void);

/* Library version string. */
const char* rca_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RCA_DAS_H */
