/*
 * rpp_audio.h -- C-ABI of the MI355X (gfx950) batched log-mel audio library.
 *
 * Drop-in for the GPU half of RPP's audio call chain: the handle API of
 * api/rpp.h and the rppt_*_gpu entry points of api/rppt_tensor_audio_augmentations.h,
 * with the same types, struct layouts, enum values and status codes as api/rppdefs.h.
 * Every entry point is unmangled (extern "C"); arguments are plain pointers and sizes.
 *
 * Memory-space rules are RPP's (api/rppt_tensor_audio_augmentations.h):
 *   src/dst tensors         device memory
 *   per-clip arrays         pinned host memory or device memory (kernels read them);
 *                           rppt_spectrogram_gpu additionally needs nothing on the host
 *   windowFunction          host memory (copied to the device, cached per content)
 *   RpptResamplingWindow    host struct; its lookupPinned table pinned or device memory
 * All work is enqueued on the handle's stream; no entry point blocks the host except
 * where noted.  Handles are not thread-safe (one handle per host thread).
 */
#ifndef RPP_AUDIO_MI355X_H
#define RPP_AUDIO_MI355X_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>   /* the reference's signatures take bool */
#endif

#ifdef __cplusplus
#include <vector>
#include <xmmintrin.h>
extern "C" {
#endif

/* ---------------- scalar typedefs (api/rppdefs.h:92-113) ---------------- */
typedef unsigned char Rpp8u;
typedef int Rpp32s;
typedef unsigned int Rpp32u;
typedef long long Rpp64s;
typedef float Rpp32f;
typedef double Rpp64f;
typedef void *RppPtr_t;
typedef size_t RppSize_t;

/* ---------------- status codes (api/rppdefs.h:117-176, :190-201) ---------------- */
typedef enum {
    RPP_SUCCESS = 0,
    RPP_ERROR = -1,
    RPP_ERROR_INVALID_ARGUMENTS = -2,
    RPP_ERROR_LOW_OFFSET = -3,
    RPP_ERROR_ZERO_DIVISION = -4,
    RPP_ERROR_HIGH_SRC_DIMENSION = -5,
    RPP_ERROR_NOT_IMPLEMENTED = -6,
    RPP_ERROR_INVALID_SRC_CHANNELS = -7,
    RPP_ERROR_INVALID_DST_CHANNELS = -8,
    RPP_ERROR_INVALID_SRC_LAYOUT = -9,
    RPP_ERROR_INVALID_DST_LAYOUT = -10,
    RPP_ERROR_INVALID_SRC_DATATYPE = -11,
    RPP_ERROR_INVALID_DST_DATATYPE = -12,
    RPP_ERROR_INVALID_SRC_OR_DST_DATATYPE = -13,
    RPP_ERROR_INSUFFICIENT_DST_BUFFER_LENGTH = -14,
    RPP_ERROR_INVALID_PARAMETER_DATATYPE = -15,
    RPP_ERROR_NOT_ENOUGH_MEMORY = -16,
    RPP_ERROR_OUT_OF_BOUND_SRC_ROI = -17,
    RPP_ERROR_LAYOUT_MISMATCH = -18,
    RPP_ERROR_INVALID_CHANNELS = -19,
    RPP_ERROR_INVALID_OUTPUT_TILE_LENGTH = -20,
    RPP_ERROR_OUT_OF_BOUND_SHARED_MEMORY_SIZE = -21,
    RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE = -22,
    RPP_ERROR_INVALID_SRC_DIMS = -23,
    RPP_ERROR_INVALID_DST_DIMS = -24,
    RPP_ERROR_INVALID_DIM_LENGTHS = -25,
    RPP_ERROR_INVALID_AXIS = -26
} RppStatus;

typedef enum {
    rppStatusSuccess = 0,
    rppStatusBadParm = -1,
    rppStatusUnknownError = -2,
    rppStatusNotInitialized = -3,
    rppStatusInvalidValue = -4,
    rppStatusAllocFailed = -5,
    rppStatusInternalError = -6,
    rppStatusNotImplemented = -7,
    rppStatusUnsupportedOp = -8
} rppStatus_t;

/* ---------------- enums (api/rppdefs.h:178-186, :398-419, :464-478) ---------------- */
typedef enum { RPP_HOST_BACKEND, RPP_HIP_BACKEND, RPP_OCL_BACKEND } RppBackend;
typedef enum { U8, F32, F16, I8, I16 } RpptDataType;
typedef enum { NCHW, NHWC, NCDHW, NDHWC, NHW, NFT, NTF } RpptLayout;
typedef enum { ZERO = 0, CLAMP, REFLECT } RpptAudioBorderType;
typedef enum { SLANEY = 0, HTK } RpptMelScaleFormula;

/* ---------------- descriptors (api/rppdefs.h:570-589, :645-649) ---------------- */
typedef struct {
    Rpp32u nStride;
    Rpp32u cStride;
    Rpp32u hStride;
    Rpp32u wStride;
} RpptStrides;

typedef struct {
    RppSize_t numDims;
    Rpp32u offsetInBytes;
    RpptDataType dataType;
    Rpp32u n, c, h, w;
    RpptStrides strides;
    RpptLayout layout;
} RpptDesc, *RpptDescPtr;

typedef struct {
    Rpp32u width;
    Rpp32u height;
} RpptImagePatch, *RpptImagePatchPtr;

/* opaque handle (api/rpp.h:64-78) */
typedef struct rppHandle *rppHandle_t;
typedef rppHandle_t RppHandle_t;

#ifdef __cplusplus
}  /* extern "C" */

/* Resampling window (api/rppdefs.h:751-790).  Same member order and layout (96 bytes on
 * x86-64); the GPU entry point reads scale, center, lobes, lookupSize and lookupPinned. */
typedef struct RpptResamplingWindow {
    Rpp32f scale = 1, center = 1;
    Rpp32s lobes = 0, coeffs = 0;
    Rpp32s lookupSize = 0;
    Rpp32f *lookupPinned = nullptr;
    std::vector<Rpp32f> lookup;
    __m128 pCenter, pScale;
} RpptResamplingWindow;

extern "C" {
#endif

/* ======================= handle API (api/rpp.h:132-244) ======================= */

/* rppCreate: api/rpp.h:132, src/modules/handle_api.cpp:29-44.  RPP_HIP_BACKEND handles serve the
 * rppt_*_gpu entry points on the current device (stream NULL = null stream, else a hipStream_t);
 * RPP_HOST_BACKEND handles serve the rppt_*_host entry points and touch no GPU.  numThreads: host
 * threads of the _host entry points, at most the hardware threads, 0 = one per clip of the batch
 * (src/modules/handle_host.cpp:54-62). */
rppStatus_t rppCreate(rppHandle_t *handle, size_t nBatchSize, Rpp32u numThreads, void *stream, RppBackend backend);
/* rppDestroy: api/rpp.h:150, handle_api.cpp:46-60 */
rppStatus_t rppDestroy(rppHandle_t handle, RppBackend backend);
/* rppSetBatchSize / rppGetBatchSize: api/rpp.h:168/:186, handle_api.cpp:62-72 */
rppStatus_t rppSetBatchSize(rppHandle_t handle, size_t batchSize);
rppStatus_t rppGetBatchSize(rppHandle_t handle, size_t *batchSize);
/* rppSetStream / rppGetStream: api/rpp.h:206/:224, handle_api.cpp:74-83 (hipStream_t) */
rppStatus_t rppSetStream(rppHandle_t handle, void *streamID);
rppStatus_t rppGetStream(rppHandle_t handle, void **streamID);

/* ======================= audio entry points ======================= */

/* api/rppt_tensor_audio_augmentations.h:156; dispatch rppt_tensor_audio_augmentations.cpp:452-484.
 * y[0] = x[0] - c*border, y[t] = x[t] - c*x[t-1] over [0, srcLengthTensor[b]). */
RppStatus rppt_pre_emphasis_filter_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                       RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor, Rpp32f *coeffTensor,
                                       RpptAudioBorderType borderType, rppHandle_t rppHandle);

/* api/rppt_tensor_audio_augmentations.h:232; dispatch rppt_tensor_audio_augmentations.cpp:579-629.
 * Power (power == 2: |X|^2, else |X|) spectrogram of centered/reflect-padded Hann frames,
 * NFT (dst[k*hStride + f]) or NTF (dst[f*hStride + k]).  FP32 arithmetic. */
RppStatus rppt_spectrogram_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                               Rpp32s *srcLengthTensor, bool centerWindows, bool reflectPadding,
                               Rpp32f *windowFunction, Rpp32s nfft, Rpp32s power, Rpp32s windowLength,
                               Rpp32s windowStep, rppHandle_t rppHandle);

/* api/rppt_tensor_audio_augmentations.h:274; dispatch rppt_tensor_audio_augmentations.cpp:488-535.
 * srcDims[2b] = numBins, srcDims[2b+1] = frames; NFT in, NFT out. */
RppStatus rppt_mel_filter_bank_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                   Rpp32s *srcDims, Rpp32f maxFreq, Rpp32f minFreq, RpptMelScaleFormula melFormula,
                                   Rpp32s numFilter, Rpp32f sampleRate, bool normalize, rppHandle_t rppHandle);

/* api/rppt_tensor_audio_augmentations.h:122; dispatch rppt_tensor_audio_augmentations.cpp:373-412.
 * out = multiplier*log10(max(minRatio, x/ref)); ref = referenceMagnitude or, if 0, the clip max. */
RppStatus rppt_to_decibels_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                               RpptImagePatchPtr srcDims, Rpp32f cutOffDB, Rpp32f multiplier,
                               Rpp32f referenceMagnitude, rppHandle_t rppHandle);

/* api/rppt_tensor_audio_augmentations.h:188; dispatch rppt_tensor_audio_augmentations.cpp:416-448 (HOST rules :156-186).
 * srcDimsTensor[2b] = samples, [2b+1] = channels (interleaved input). */
RppStatus rppt_down_mixing_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                               Rpp32s *srcDimsTensor, bool normalizeWeights, rppHandle_t rppHandle);

/* api/rppt_tensor_audio_augmentations.h:86; dispatch rppt_tensor_audio_augmentations.cpp:332-369.
 * Begin and length of each clip's non-silent region from the moving mean square of the HOST
 * kernel (src/modules/tensor/cpu/kernel/non_silent_region_detection.cpp:93-191), evaluated in
 * its order.  srcLengthTensor and both outputs in device or pinned memory; asynchronous. */
RppStatus rppt_non_silent_region_detection_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, Rpp32s *srcLengthTensor,
                                               Rpp32s *detectedIndexTensor, Rpp32s *detectionLengthTensor,
                                               Rpp32f cutOffDB, Rpp32s windowLength, Rpp32f referencePower,
                                               Rpp32s resetInterval, rppHandle_t rppHandle);

#ifdef __cplusplus
/* api/rppt_tensor_audio_augmentations.h:310; dispatch rppt_tensor_audio_augmentations.cpp:539-575.
 * Windowed-sinc resampling with the caller's lookup table. */
RppStatus rppt_resample_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                            Rpp32f *inRateTensor, Rpp32f *outRateTensor, Rpp32s *srcDimsTensor,
                            RpptResamplingWindow &window, rppHandle_t rppHandle);
#endif

/* ======================= HOST backend (buffers in host memory) ======================= */
/* The rppt_*_host entry points of api/rppt_tensor_audio_augmentations.h on the CPU cores, with a
 * RPP_HOST_BACKEND handle; every pointer is host memory.  Pre-emphasis, mel filter bank,
 * to_decibels, down-mixing, resampling and region detection give RPP HOST's outputs bit for bit;
 * the spectrogram uses this library's own FFT (RPP links FFTS), so it agrees to FP32 rounding. */

/* api/rppt_tensor_audio_augmentations.h:139; cpu/kernel/pre_emphasis_filter.cpp:27-75 */
RppStatus rppt_pre_emphasis_filter_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                        RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor, Rpp32f *coeffTensor,
                                        RpptAudioBorderType borderType, rppHandle_t rppHandle);
/* api/rppt_tensor_audio_augmentations.h:210; cpu/kernel/spectrogram.cpp:71-251 (power == 1: |X|,
 * else |X|^2, the HOST rule) */
RppStatus rppt_spectrogram_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                Rpp32s *srcLengthTensor, bool centerWindows, bool reflectPadding,
                                Rpp32f *windowFunction, Rpp32s nfft, Rpp32s power, Rpp32s windowLength,
                                Rpp32s windowStep, rppHandle_t rppHandle);
/* api/rppt_tensor_audio_augmentations.h:253; cpu/kernel/mel_filter_bank.cpp:27-198 */
RppStatus rppt_mel_filter_bank_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                    Rpp32s *srcDims, Rpp32f maxFreq, Rpp32f minFreq, RpptMelScaleFormula melFormula,
                                    Rpp32s numFilter, Rpp32f sampleRate, bool normalize, rppHandle_t rppHandle);
/* api/rppt_tensor_audio_augmentations.h:104; cpu/kernel/to_decibels.cpp:27-99 */
RppStatus rppt_to_decibels_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                RpptImagePatchPtr srcDims, Rpp32f cutOffDB, Rpp32f multiplier,
                                Rpp32f referenceMagnitude, rppHandle_t rppHandle);
/* api/rppt_tensor_audio_augmentations.h:172; cpu/kernel/down_mixing.cpp:28-122 */
RppStatus rppt_down_mixing_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                Rpp32s *srcDimsTensor, bool normalizeWeights, rppHandle_t rppHandle);
/* api/rppt_tensor_audio_augmentations.h:65; cpu/kernel/non_silent_region_detection.cpp:93-191 */
RppStatus rppt_non_silent_region_detection_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, Rpp32s *srcLengthTensor,
                                                Rpp32s *detectedIndexTensor, Rpp32s *detectionLengthTensor,
                                                Rpp32f cutOffDB, Rpp32s windowLength, Rpp32f referencePower,
                                                Rpp32s resetInterval, rppHandle_t rppHandle);
#ifdef __cplusplus
/* api/rppt_tensor_audio_augmentations.h:292; cpu/kernel/resample.cpp:27-146 (reads window.lookup,
 * window.lookupSize, scale, center, lobes) */
RppStatus rppt_resample_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                             Rpp32f *inRateTensor, Rpp32f *outRateTensor, Rpp32s *srcDimsTensor,
                             RpptResamplingWindow &window, rppHandle_t rppHandle);
#endif

/* ======================= extensions (no counterpart in the reference) ======================= */

#ifdef __cplusplus
/* rppt_down_mixing_gpu followed by rppt_resample_gpu in one launch (the down-mixed signal stays
 * on chip): src interleaved [n][samples * channels] with srcDimsTensor[2b] = samples,
 * [2b+1] = channels; dst mono [n][nStride].  Bit-identical to the two calls in sequence. */
RppStatus rppt_down_mix_resample_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                     Rpp32f *inRateTensor, Rpp32f *outRateTensor, Rpp32s *srcDimsTensor,
                                     bool normalizeWeights, RpptResamplingWindow &window, rppHandle_t rppHandle);
#endif

/* Arithmetic of the fused DFT / mel contraction. */
typedef enum {
    RPPT_LOGMEL_FP32 = 0,      /* FP32 vector FMA throughout */
    RPPT_LOGMEL_BF16X3 = 1,    /* MFMA bf16, split hi+lo: hi*hi + hi*lo + lo*hi, FP32 accumulate */
    RPPT_LOGMEL_FP16X3 = 2,    /* MFMA fp16, split hi+lo with power-of-two scaling, FP32 accumulate */
    RPPT_LOGMEL_BF16 = 3,      /* MFMA bf16, single product (fast, reported error only) */
    RPPT_LOGMEL_BF16X6 = 4     /* MFMA bf16, three-term split hi+mid+lo of both DFT operands: the six
                                  products down to 2^-16 relative, FP32 accumulate */
} RpptLogMelPrecision;

/* Parameters of the whole log-mel chain, the union of the four reference calls' arguments. */
typedef struct {
    Rpp32f preEmphasisCoeff;        /* rppt_pre_emphasis_filter coeff (same for every clip) */
    RpptAudioBorderType borderType; /* pre-emphasis border */
    Rpp32s nfft, windowLength, windowStep;
    Rpp32s centerWindows, reflectPadding;
    Rpp32s power;                   /* 2: power, else magnitude */
    Rpp32f sampleRate, minFreq, maxFreq;
    Rpp32s numFilter;
    RpptMelScaleFormula melFormula;
    Rpp32s normalize;
    Rpp32f cutOffDB, multiplier, referenceMagnitude;  /* referenceMagnitude >= 0; 0 = each clip's maximum */
} RpptLogMelParams;

/* Whole chain pre-emphasis -> spectrogram -> mel -> to_decibels over a ragged batch in one
 * persistent launch.  src [n][nStride] f32 (device), srcLengthTensor device or pinned.
 * dst NFT [n][numFilter >= h rows][hStride] f32 (a clip's frames beyond hStride are not written)
 * or NTF [n][h][hStride >= numFilter] (frames beyond h are not written).  Equivalent to the four
 * rppt_*_gpu calls with the matching arguments.  referenceMagnitude 0 (each clip's maximum, as
 * rppt_to_decibels_gpu) adds a second in-place pass over the output after the fused launch. */
RppStatus rppt_logmel_fused_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                Rpp32s *srcLengthTensor, const RpptLogMelParams *params,
                                RpptLogMelPrecision precision, rppHandle_t rppHandle);

/* The fused launch over a region of each row: samples [srcBeginTensor[b], + srcLengthTensor[b])
 * (begin clamped to the row, the length to what remains of it), e.g. the begin/length pair
 * rppt_non_silent_region_detection_gpu writes.  Same output as rppt_logmel_fused_gpu on the
 * trimmed clips, with no copy of them.  Both arrays device or pinned. */
RppStatus rppt_logmel_fused_region_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                       RpptDescPtr dstDescPtr, Rpp32s *srcBeginTensor, Rpp32s *srcLengthTensor,
                                       const RpptLogMelParams *params, RpptLogMelPrecision precision,
                                       rppHandle_t rppHandle);

#ifdef __cplusplus
/* Down-mix, resample and the fused log-mel chain: src interleaved [n][samples * channels] f32 at
 * inRate (srcDimsTensor[2b] = samples per channel, [2b+1] = channels; device or pinned), dst as
 * rppt_logmel_fused_gpu at outRate, each clip's length ceil(samples * outRate / inRate) -- the output
 * of rppt_down_mix_resample_gpu (one rate pair for the batch) fed to rppt_logmel_fused_gpu, bit for
 * bit.  oneLaunch != 0 runs it as one launch whose work items compute the resampled samples they
 * read (the resampled clips never reach memory) where that launch applies: FP16X3, nfft 512, hop
 * 160, power 2, NFT output, referenceMagnitude != 0, inRate / outRate = 2 or 3 and window.lobes = 16;
 * otherwise, and with oneLaunch = 0, the two launches through a buffer of the handle. */
RppStatus rppt_logmel_front_end_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                    Rpp32s *srcDimsTensor, Rpp32f inRate, Rpp32f outRate, bool normalizeWeights,
                                    RpptResamplingWindow &window, const RpptLogMelParams *params,
                                    RpptLogMelPrecision precision, Rpp32s oneLaunch, rppHandle_t rppHandle);
#endif

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * output as rppt_logmel_fused_gpu, but the call never allocates: the DFT and mel tables, the tile
 * plan, tile records and per-clip maxima live in one 16 MiB device array of the library (one copy
 * per device, held until the process ends).  The tables of a parameter set are built there on the
 * first call with it (that call must not be under stream capture); a call with other parameters
 * rebuilds them in place.  RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE when a batch's plan does not
 * fit what the tables leave.  Every call on a device shares that array: calls that overlap on one
 * device (other streams, other threads, other handles) race. */
RppStatus rppt_logmel_fused_noalloc_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                        RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                        const RpptLogMelParams *params, RpptLogMelPrecision precision,
                                        rppHandle_t rppHandle);

/* Packed variant for serving: clip b's samples at srcPtr + srcOffsets[b] (srcLengthTensor[b] of
 * them), its dB output as a dense [numFilter][frames_b] block at dstPtr + dstOffsets[b], with
 * frames_b = L/H + 1 (centered) or (L-W)/H + 1.  Offsets and lengths in device or pinned memory. */
RppStatus rppt_logmel_fused_packed_gpu(const Rpp32f *srcPtr, const long long *srcOffsets, const Rpp32s *srcLengthTensor,
                                       Rpp32s batchSize, Rpp32f *dstPtr, const long long *dstOffsets,
                                       const RpptLogMelParams *params, RpptLogMelPrecision precision,
                                       rppHandle_t rppHandle);

/* The four drop-in calls behind one entry point, intermediates in the handle's workspace:
 * perClip != 0 issues them once per clip (one launch per clip per stage), else once for the batch.
 * srcLengthTensor must be host-readable (host or pinned memory). */
RppStatus rppt_logmel_chain_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                                Rpp32s *srcLengthTensor, const RpptLogMelParams *params, int perClip,
                                rppHandle_t rppHandle);

/* Launch shape of the fused kernel; 0 fields mean "default". */
typedef struct {
    Rpp32s framesPerTile;   /* 16 selects the factored-DFT kernel where it applies (fp16x3; nfft 512
                               with hop 160, 16 frames per wave item; nfft 1024 with hop 256, radix 2,
                               8 frames per wave item, power 2, NFT, referenceMagnitude != 0; a filter
                               bank whose 32-bin chunks each meet at most six 16-filter tiles, or any
                               bank of at most 80 filters); the
                               dense MFMA kernel runs 64 frames (nfft <= 512; at nfft 512, 32 on
                               request) or 32 (above); the FP32 kernel 32; 0 = default (factored
                               where it applies) */
    Rpp32s blocksPerCU;     /* persistent workgroups per compute unit (dense kernels; default 2) */
} RpptLogMelTiles;
RppStatus rppt_logmel_set_tiles(rppHandle_t rppHandle, const RpptLogMelTiles *tiles);
/* Autotuner: times every supported tile shape on this batch (`iters` launches each, events on the
 * handle's stream), keeps the fastest in the handle and reports it. */
RppStatus rppt_logmel_autotune(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                               Rpp32s *srcLengthTensor, const RpptLogMelParams *params,
                               RpptLogMelPrecision precision, Rpp32s iters, rppHandle_t rppHandle,
                               RpptLogMelTiles *chosen, Rpp32f *bestMs);
RppStatus rppt_logmel_get_tiles(rppHandle_t rppHandle, RpptLogMelTiles *tiles);

/* ======================= serving (no counterpart in the reference) ======================= */
/* A continuous-batching request loop over rppt_logmel_fused_packed_gpu on the current device:
 * host threads submit clips, a scheduler thread batches them onto a pool of device slots (each a
 * stream with device-resident buffers) and stamps each request's completion time.  A submitted clip
 * is copied once, into a pinned arena the batch's host-to-device copy reads (a clip that finds the
 * arena full is held in its own buffer and staged instead); a clip larger than the slots were sized
 * for is admitted when the free device memory at that moment (hipMemGetInfo) can hold a slot for
 * it, and an idle slot is regrown for it. */
typedef struct rppaServer rppaServer;
typedef struct {
    int maxBatch;             /* clips per batch (upper bound) */
    int deadlineUs;           /* longest a clip waits for its batch to close while a slot is free */
    long long memBudgetBytes; /* device bytes for all slots (0: 80% of free HBM, at most 8 GiB) */
    int numSlots;             /* concurrent batches (streams), 1..16 */
    int eager;                /* close a batch as soon as a slot is idle */
    int keepOutputs;          /* keep dB rows for rppa_server_wait (else timings only) */
    int growSlots;            /* admit a clip larger than a slot when 80% of the free device memory
                                 at submit covers a slot grown for it (else refuse it) */
} rppaServeConfig;
typedef struct {
    long long submitted, completed, refused, batches;
    double p50Us, p90Us, p99Us, meanUs, maxUs;   /* arrival -> result in host memory */
    double meanBatchClips, audioSamples, gpuBusySeconds;
    long long slotCapacitySamples;
    double maxQueueUs, maxServiceUs;   /* the slowest request: arrival -> its batch launched, launch -> done */
    long long maxLatencySamples;       /* and its length */
    double maxBatchGpuUs;              /* its batch's copies and launch on the device (events) */
    double maxLoopUs, maxLaunchUs;     /* the scheduler's longest loop turn, and longest batch launch call */
    double maxCopyInCallUs, maxKernelCallUs, maxCopyOutCallUs;   /* longest host time of each part of it */
    long long arenaSubmits, stagedSubmits, slotGrowths;
} rppaServeStats;
RppStatus rppa_server_create(const RpptLogMelParams *params, RpptLogMelPrecision precision,
                             const rppaServeConfig *config, rppaServer **server);
RppStatus rppa_server_destroy(rppaServer *server);
/* copies `length` host samples; RPP_ERROR_NOT_ENOUGH_MEMORY when no slot can hold the clip (nor be
 * grown for it within the free device memory) */
RppStatus rppa_server_submit(rppaServer *server, const float *samples, int length, long long *id);
/* blocks until `id` is done; out receives [numFilter][frames] dB rows when outputs are kept.  Several
 * threads may wait on one id (the request is freed when the last returns).  Without kept outputs a
 * request that already finished reports success with *frames = -1 (out must then be NULL). */
RppStatus rppa_server_wait(rppaServer *server, long long id, float *out, long long outCapacity, int *frames,
                           double *latencyUs);
RppStatus rppa_server_drain(rppaServer *server, int timeoutMs);
RppStatus rppa_server_stats(rppaServer *server, rppaServeStats *stats);

#ifdef __cplusplus
}
#endif

#endif /* RPP_AUDIO_MI355X_H */
