/*
 * libmetalmom.so -- MI355X (gfx950) implementation of MetalMom's spectral
 * front end: STFT -> power -> mel filterbank -> power_to_db -> DCT-II (MFCC).
 *
 * Part 1 is a drop-in for the reference C ABI: the same status codes, struct
 * layouts and prototypes as Sources/MetalMomCBridge/include/metalmom.h of
 * zakkeown/MetalMom, for the hot-path exports. Part 2 is additive: complex
 * STFT, device-resident batches (one persistent launch per step, optionally
 * replayed from a captured hipGraph), an autotuner and a request server.
 */
#ifndef METALMOM_H
#define METALMOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Part 1: drop-in ABI
 * ---------------------------------------------------------------------- */

/* Status codes -- reference metalmom.h:7-11. MM_ERR_METAL_UNAVAILABLE keeps
 * its value and means "no usable GPU" here. */
#define MM_OK 0
#define MM_ERR_INVALID_INPUT -1
#define MM_ERR_METAL_UNAVAILABLE -2
#define MM_ERR_ALLOC_FAILED -3
#define MM_ERR_INTERNAL -4

/* Output buffer -- reference metalmom.h:14-20. `data` is malloc'd by the
 * library and released by mm_buffer_free. dtype 0 = float32; dtype 1 =
 * complex64 stored as interleaved (re, im) floats, `count` complex elements. */
typedef struct {
    float* data;
    int64_t shape[8];
    int32_t ndim;
    int32_t dtype;
    int64_t count;
} MMBuffer;

/* STFT parameters -- reference metalmom.h:23-28. */
typedef struct {
    int32_t n_fft;
    int32_t hop_length;
    int32_t win_length;
    int32_t center;
} MMSTFTParams;

/* Opaque context -- reference metalmom.h:31. Not thread-safe: one per
 * thread. Contexts on different threads run concurrently (each has its own
 * HIP stream and workspace). */
typedef void* mm_context;

/* Replaces Bridge.swift:22-30 (mm_init). Cheap; the GPU is opened lazily. */
mm_context mm_init(void);
/* Replaces Bridge.swift:33-40 (mm_destroy). A second destroy is a no-op. */
void mm_destroy(mm_context ctx);

/* Replaces Bridge.swift:44-140 (mm_stft): magnitude, row-major
 * [n_fft/2+1, n_frames], zero centre padding. n_fft must be a power of two. */
int32_t mm_stft(mm_context ctx, const float* signal_data, int64_t signal_length,
                int32_t sample_rate, const MMSTFTParams* params, MMBuffer* out);

/* Replaces Bridge.swift:144-221 (mm_istft), reference metalmom.h:41-45:
 * inverse STFT of a complex spectrogram given as interleaved (re, im) floats,
 * row-major [n_freqs, n_frames] (stft_count = 2 n_freqs n_frames). n_fft =
 * 2 (n_freqs - 1); windowed overlap-add divided by the summed squared window
 * where it exceeds 1e-10 (STFT.swift:686-861). center trims n_fft/2 at both
 * ends; output_length > 0 cuts or zero-pads the result to that length; a
 * non-power-of-two n_fft gives n_fft + (n_frames-1) hop zeros. 1-D output.
 * Rejected (-1) where the reference would trap or this library has no kernel:
 * a stft_count that does not match the shape, hop < 0, a power-of-two n_fft
 * above 16384, or 2^31 output samples or more. */
int32_t mm_istft(mm_context ctx, const float* stft_data, int64_t stft_count,
                 int32_t n_freqs, int32_t n_frames, int32_t sample_rate,
                 int32_t hop_length, int32_t win_length, int32_t center,
                 int64_t output_length, MMBuffer* out);

/* Replaces Bridge.swift:1319-1358 (mm_onset_strength), reference
 * metalmom.h:170-173: onset strength envelope (OnsetDetection.swift:23-128):
 * power mel spectrogram (win = n_fft), power_to_db(ref 1, amin 1e-10,
 * top_db 80) over the whole matrix, lag-1 half-wave rectified difference,
 * mean over the bands (aggregate != 0: [1, frames]) or per band
 * ([n_mels, frames]), 1 + n_fft/(2 hop) leading zeros with center (1
 * without), cut to the frame count with center. hop <= 0 -> 512, f_max <= 0
 * -> sr/2. A clip of at most one frame (or a non-power-of-two n_fft) gives
 * one zero column. ctx may be NULL (the reference does not check it).
 * Rejected (-1) where the reference would divide by zero or produce NaN:
 * sample_rate <= 0, n_mels < 1. */
int32_t mm_onset_strength(mm_context ctx, const float* signal_data, int64_t signal_length,
                          int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                          int32_t n_mels, float f_min, float f_max,
                          int32_t center, int32_t aggregate, MMBuffer* out);

/* Replaces Bridge.swift:225-263 (mm_amplitude_to_db): 20 log10(max(x, amin)/ref),
 * clipped to max - top_db when top_db > 0. 1-D output of `count` values. */
int32_t mm_amplitude_to_db(mm_context ctx, const float* data, int64_t count,
                           float ref, float amin, float top_db, MMBuffer* out);

/* Replaces Bridge.swift:267-306 (mm_power_to_db): 10 log10(max(x, amin)/ref),
 * clipped to max - top_db when top_db > 0. 1-D output of `count` values. */
int32_t mm_power_to_db(mm_context ctx, const float* data, int64_t count,
                       float ref, float amin, float top_db, MMBuffer* out);

/* Replaces Bridge.swift:310-400 (mm_mel_spectrogram): [n_mels, n_frames].
 * hop <= 0 -> n_fft/4, win <= 0 -> n_fft, f_max <= 0 -> sr/2. */
int32_t mm_mel_spectrogram(mm_context ctx, const float* signal_data, int64_t signal_length,
                           int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                           int32_t win_length, int32_t center, float power,
                           int32_t n_mels, float f_min, float f_max,
                           MMBuffer* out);

/* Replaces Bridge.swift:494-545 (mm_chroma_stft), reference metalmom.h:69-72:
 * chroma filterbank (Chroma.swift:158-275: Gaussian pitch-class bumps, L2
 * column norm, octave weighting around octave 5, rows starting at C, tuning
 * 0) times the power spectrogram, [n_chroma, n_frames], no frame
 * normalisation. hop <= 0 -> n_fft/4, win <= 0 -> n_fft; a non-power-of-two
 * n_fft gives [n_chroma, 0]. Rejected (-1): n_chroma outside 1..256 (no
 * kernel), sample_rate <= 0. */
int32_t mm_chroma_stft(mm_context ctx, const float* signal_data, int64_t signal_length,
                       int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                       int32_t win_length, int32_t n_chroma, int32_t center,
                       MMBuffer* out);

/* Replace Bridge.swift:621-813 (mm_spectral_centroid / _bandwidth / _rolloff
 * / _flatness), reference metalmom.h:76-90: per-frame reductions of the
 * magnitude STFT (SpectralDescriptors.swift), 1-D [n_frames]. Frequencies
 * are Float(k) * sr / n_fft. centroid: sum f S / sum S (0 without energy);
 * bandwidth: (sum S |f - centroid|^p / sum S)^(1/p); rolloff: first bin
 * whose cumulative magnitude reaches roll_percent of the total; flatness:
 * geometric / arithmetic mean of max(S^2, 1e-10) (0 when sum S < 1e-10).
 * hop <= 0 -> n_fft/4, win <= 0 -> n_fft; a non-power-of-two n_fft (no STFT
 * frames) gives an empty result. */
int32_t mm_spectral_centroid(mm_context ctx, const float* signal_data, int64_t signal_length,
                             int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                             int32_t win_length, int32_t center, MMBuffer* out);
int32_t mm_spectral_bandwidth(mm_context ctx, const float* signal_data, int64_t signal_length,
                              int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                              int32_t win_length, int32_t center, float p, MMBuffer* out);
/* Replaces Bridge.swift:698-737 (mm_spectral_contrast), reference
 * metalmom.h:82-84: per frame and band (bins [0, fmin), then octaves from
 * f_min, the last up to sr/2; SpectralDescriptors.swift:137-204),
 * 10 log10(mean of the top q magnitudes) - 10 log10(mean of the bottom q),
 * q = max(1, round(0.02 * bins)), each floored at 1e-10; an empty band gives
 * 0. [n_bands + 1, n_frames]. hop <= 0 -> n_fft/4, win <= 0 -> n_fft; a
 * non-power-of-two n_fft gives [n_bands + 1, 0]. Rejected (-1) where the
 * reference traps or divides by zero: n_bands < 1, sample_rate <= 0; and
 * n_bands > 62. */
int32_t mm_spectral_contrast(mm_context ctx, const float* signal_data, int64_t signal_length,
                             int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                             int32_t win_length, int32_t center, int32_t n_bands, float f_min,
                             MMBuffer* out);
int32_t mm_spectral_rolloff(mm_context ctx, const float* signal_data, int64_t signal_length,
                            int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                            int32_t win_length, int32_t center, float roll_percent, MMBuffer* out);
int32_t mm_spectral_flatness(mm_context ctx, const float* signal_data, int64_t signal_length,
                             int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                             int32_t win_length, int32_t center, MMBuffer* out);

/* Replace Bridge.swift:1008-1085 (mm_get_duration, mm_get_sample_rate,
 * mm_load), reference metalmom.h:125-131. The reference decodes through
 * AVFoundation; this library reads RIFF/WAVE files: PCM u8 / s16 / s24 / s32
 * (also WAVE_FORMAT_EXTENSIBLE) and IEEE float32 / float64, any channel
 * count; other files give MM_ERR_INVALID_INPUT. mm_load (AudioIO.swift:16-131):
 * sr <= 0 keeps the file's rate, start = trunc(offset * file rate), duration
 * > 0 reads at most trunc(duration * file rate) frames, no frames left gives
 * an empty signal; 1-D float32 output (integer PCM scaled by 2^-(bits-1));
 * channel 0 of a multichannel file, or the mean of its channels when mono !=
 * 0; another rate is reached with mm_resample's converter on the GPU (the
 * reference uses AVAudioConverter); *out_sr (may be NULL) receives the output
 * rate. ctx may be NULL. A negative offset is rejected (-1). */
int32_t mm_get_duration(mm_context ctx, const char* path, float* out_duration);
int32_t mm_get_sample_rate(mm_context ctx, const char* path, int32_t* out_sr);
int32_t mm_load(mm_context ctx, const char* path, int32_t sr, int32_t mono, float offset,
                float duration, MMBuffer* out, int32_t* out_sr);

/* Replaces Bridge.swift:1087-1104 (mm_resample), reference metalmom.h:134-135:
 * windowed-sinc sample-rate conversion (Resample.swift:12-96; sinc x Kaiser
 * beta 6, 64 zero crossings, 512-point table with linear interpolation, the
 * filter widened by target/source when downsampling), 1-D output of
 * ceil(signal_length * target_sr / source_sr) samples; equal rates return a
 * copy of the input. ctx may be NULL (the reference does not check it).
 * Rejected (-1) where the reference would divide by zero or trap:
 * source_sr <= 0 or target_sr <= 0. */
int32_t mm_resample(mm_context ctx, const float* signal_data, int64_t signal_length,
                    int32_t source_sr, int32_t target_sr, MMBuffer* out);

/* Replaces Bridge.swift:402-493 (mm_mfcc): [min(n_mfcc, n_mels), n_frames];
 * power 2 mel, power_to_db(ref 1, amin 1e-10, top_db 80), DCT-II ortho. */
int32_t mm_mfcc(mm_context ctx, const float* signal_data, int64_t signal_length,
                int32_t sample_rate, int32_t n_mfcc, int32_t n_fft,
                int32_t hop_length, int32_t win_length,
                int32_t n_mels, float f_min, float f_max,
                int32_t center, MMBuffer* out);

/* Replaces Bridge.swift:3695-3711 (mm_buffer_free). Null-safe; zeroes *buf. */
void mm_buffer_free(MMBuffer* buf);

/* ------------------------------------------------------------------------
 * Part 2: additive API (no counterpart in the reference ABI)
 * ---------------------------------------------------------------------- */

/* Output kinds and precisions. */
#define MM_KIND_STFT_MAG 0
#define MM_KIND_STFT_POWER 1
#define MM_KIND_STFT_COMPLEX 2
#define MM_KIND_MEL 3
#define MM_KIND_MFCC 4
#define MM_KIND_CHROMA 5 /* STFT chroma (mm_chroma_stft): n_mels = n_chroma (1..256), f_min / f_max unused */

#define MM_PREC_F32 0   /* FP32 mel/DCT contraction (FP32 MFMA) */
#define MM_PREC_BF16 1  /* BF16 MFMA, FP32 accumulation */
#define MM_PREC_F16 2   /* FP16 MFMA, FP32 accumulation */

#define MM_INPUT_F32 0
#define MM_INPUT_F16 1

#define MM_PAD_CONSTANT 0 /* zeros: the reference's centre padding */
#define MM_PAD_REFLECT 1  /* numpy.pad(mode="reflect") */

/* Full request description. Sentinels follow the drop-in calls: hop <= 0 ->
 * n_fft/4, win <= 0 -> n_fft, f_max <= 0 -> sr/2, top_db <= 0 -> no clamp. */
typedef struct {
    int32_t kind;
    int32_t sample_rate;
    int32_t n_fft;
    int32_t hop_length;
    int32_t win_length;
    int32_t center;
    int32_t pad_mode;
    float power;
    int32_t n_mels;
    float f_min;
    float f_max;
    int32_t n_mfcc;
    float top_db;
    int32_t precision;
    int32_t input_dtype;
} MMSpecParams;

/* Complex64 STFT of one clip, [n_fft/2+1, n_frames], dtype 1 (see MMBuffer).
 * Same framing as mm_stft (STFT.swift:512-668 computeComplex, not exported by
 * the reference). */
int32_t mm_stft_complex(mm_context ctx, const float* signal_data, int64_t signal_length,
                        int32_t sample_rate, const MMSTFTParams* params, MMBuffer* out);

/* Spectral descriptor kinds of mm_spectral_descriptor_device (param:
 * bandwidth p, rolloff roll_percent; unused otherwise). */
#define MM_DESC_CENTROID 0
#define MM_DESC_BANDWIDTH 1
#define MM_DESC_ROLLOFF 2
#define MM_DESC_FLATNESS 3
/* The mm_spectral_* descriptors on device memory: d_signal holds
 * signal_length samples, d_out receives one value per STFT frame (count
 * through out_frames, which may be NULL). Enqueued on `stream` without a host
 * synchronisation after the first call for a clip length and parameter set.
 * Needs a live context (it holds the magnitude spectrogram in between); a
 * later call on the same context from another stream (or a host-buffer call)
 * first waits for this call's work. A non-power-of-two n_fft is rejected. */
int32_t mm_spectral_descriptor_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                                      int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                                      int32_t win_length, int32_t center, int32_t kind, float param,
                                      float* d_out, int64_t* out_frames, void* stream);
/* mm_spectral_contrast on device memory: d_out receives [n_bands + 1,
 * frames] (frames through out_frames, which may be NULL). Enqueued on
 * `stream` without a host synchronisation after the first call for a clip
 * length and parameter set. Needs a live context; a later call on the same
 * context from another stream (or a host-buffer call) first waits for this
 * call's work. A non-power-of-two n_fft is rejected. */
int32_t mm_spectral_contrast_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                                    int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                                    int32_t win_length, int32_t center, int32_t n_bands, float f_min,
                                    float* d_out, int64_t* out_frames, void* stream);
/* mm_onset_strength on device memory: d_signal holds signal_length samples,
 * d_out receives the [rows, cols] result (rows 1 or n_mels, cols frames;
 * both reported through out_rows / out_cols, which may be NULL; size d_out
 * for n_mels x (frames of the clip) floats when unsure). Enqueued on `stream`
 * without a host synchronisation after the first call for a clip length and
 * parameter set (which uploads the clip's work list and, on first use of a
 * configuration, builds its tables). Needs a live context (it holds the mel
 * spectrogram in between); a later call on the same context from another
 * stream (or a host-buffer call) first waits for this call's work. A
 * non-power-of-two n_fft is rejected (-1). */
int32_t mm_onset_strength_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                                 int32_t sample_rate, int32_t n_fft, int32_t hop_length,
                                 int32_t n_mels, float f_min, float f_max, int32_t center,
                                 int32_t aggregate, float* d_out, int64_t* out_rows,
                                 int64_t* out_cols, void* stream);

/* Number of samples mm_resample returns (ceil(n * target / source) in
 * double; n for equal rates), or -1 for a rate <= 0 or n < 0. */
int64_t mm_resample_length(int64_t signal_length, int32_t source_sr, int32_t target_sr);
/* mm_resample on device memory: d_signal holds signal_length samples, d_out
 * mm_resample_length(...) floats (it must not overlap d_signal); enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) without a host
 * synchronisation after the first call, which builds the filter table. ctx
 * may be NULL. */
int32_t mm_resample_device(mm_context ctx, const float* d_signal, int64_t signal_length,
                           int32_t source_sr, int32_t target_sr, float* d_out, void* stream);

/* Number of samples mm_istft returns for this shape, or -1 if it would
 * reject it. */
int64_t mm_istft_length(int32_t n_freqs, int32_t n_frames, int32_t hop_length, int32_t center,
                        int64_t output_length);
/* mm_istft on device memory: d_stft holds the interleaved spectrogram with
 * rows of ld_frames complex values (ld_frames >= n_frames; 0 = n_frames; a
 * multiple of 16 lets the kernel read whole cache lines), d_out
 * mm_istft_length(...) floats; enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) without a host synchronisation. The context provides the
 * scratch of the generic path; a later call on the same context from another
 * stream (or a host-buffer call) first waits for this call's work. */
int32_t mm_istft_device(mm_context ctx, const float* d_stft, int32_t ld_frames, int32_t n_freqs,
                        int32_t n_frames, int32_t sample_rate, int32_t hop_length, int32_t win_length,
                        int32_t center, int64_t output_length, float* d_out, void* stream);

/* One clip with every option of MMSpecParams (input must be float32 here). */
int32_t mm_spectral(mm_context ctx, const MMSpecParams* params, const float* signal_data,
                    int64_t signal_length, MMBuffer* out);

/* mm_spectral on device memory that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops). d_signal holds signal_length
 * float32 samples, d_out receives the [rows, frames] result (both reported
 * through rows_out / frames_out, which may be NULL); enqueued on `stream`
 * (NULL = the null stream), kernel launches only. The clip's work list and
 * top_db scratch live in one 1 MiB device array per GPU, shared by every
 * call on that GPU and keeping the last call's contents: calls that overlap
 * on one GPU race. Limits: configurations of the fused kernel only (n_fft
 * 256..4096, n_mels <= 256, float32 input), and at most ~37,000 tiles of 16
 * frames (about 600,000 frames) per clip; beyond them -1. The
 * configuration's tables must already be resident (the first mm_spectral,
 * mm_batch_create or server request of the same configuration builds them);
 * otherwise -3 and nothing is enqueued. */
int32_t mm_spectral_device_noalloc(const MMSpecParams* params, const float* d_signal, int64_t signal_length,
                                   float* d_out, int64_t* rows_out, int64_t* frames_out, void* stream);

/* Host-buffer batch: clip i -> outs[i] (each released with mm_buffer_free).
 * All clips run in one persistent launch. */
int32_t mm_spectral_batch(mm_context ctx, const MMSpecParams* params, const float* const* clips,
                          const int64_t* lengths, int32_t n_clips, MMBuffer* outs);

/* Device-resident batch plan. Clips are packed back to back in one device
 * signal buffer (sample offsets from mm_batch_clip); clip i's output is the
 * row-major [rows, frames] block at its output offset (in floats). */
typedef void* mm_batch;
int32_t mm_batch_create(mm_context ctx, const MMSpecParams* params, const int64_t* lengths,
                        int32_t n_clips, mm_batch* out);
int32_t mm_batch_info(mm_batch b, int64_t* total_input_samples, int64_t* total_output_floats,
                      int64_t* total_frames, int32_t* n_tiles);
int32_t mm_batch_clip(mm_batch b, int32_t clip, int64_t* input_offset, int64_t* output_offset,
                      int64_t* rows, int64_t* frames);
/* Enqueue one step on `hip_stream` (NULL: the context's stream). Inputs and
 * outputs are device pointers; nothing is allocated. Steps of one plan
 * alternate between two sets of top_db scratch, each step clearing the
 * other set for the next one: a plan's steps must run in the order they
 * were enqueued (one stream, or streams ordered by events). */
int32_t mm_batch_run(mm_batch b, const void* d_signal, float* d_out, void* hip_stream);
/* Same, replayed from a hipGraph captured on first use for this
 * (d_signal, d_out, stream) and scratch set (two graphs per triple). */
int32_t mm_batch_run_graph(mm_batch b, const void* d_signal, float* d_out, void* hip_stream);
/* Persistent grid: workgroups per CU (0: autotuned or default). */
int32_t mm_batch_set_grid(mm_batch b, int32_t wg_per_cu);
/* Workgroup shapes of the fused kernel: MM_SHAPE_W16 = 16 wavefronts, one
 * frame each, 16-frame tiles (n_fft <= 2048); MM_SHAPE_W8 = 8 wavefronts,
 * 8-frame tiles; MM_SHAPE_W8F2 = 8 wavefronts with two frames each, 16-frame
 * tiles (n_fft 1024 to 4096; two workgroups per CU at 1024 and 2048);
 * MM_SHAPE_W16P = 16 wavefronts pipelined (staging rows and DCT partials
 * double-buffered, one barrier per tile; BF16 / FP16 MFCC at n_fft 2048).
 * Results do not depend on the shape. MM_SHAPE_DEFAULT: W16 up to n_fft 2048
 * (W8F2 for the STFT magnitude / power at 1024 and 2048, W16P for the 16-bit
 * MFCC at 2048),
 * W8F2 at 4096 (W8 where its LDS does not fit). */
#define MM_SHAPE_DEFAULT 0
#define MM_SHAPE_W16 1
#define MM_SHAPE_W8 2
#define MM_SHAPE_W8F2 3
#define MM_SHAPE_W16P 4
/* Rebuilds the plan's tile list for `shape` (synchronises the plan's
 * stream); -1 for a shape the configuration does not offer. */
int32_t mm_batch_set_shape(mm_batch b, int32_t shape);
/* What the next step launches: shape, wavefronts per workgroup, frames per
 * tile, persistent workgroups and the contiguous tiles each one processes
 * (all 0 for configurations outside the fused kernel). Any pointer may be
 * NULL. */
int32_t mm_batch_launch_shape(mm_batch b, int32_t* shape, int32_t* waves_per_wg, int32_t* tile_frames,
                              int32_t* grid, int32_t* tiles_per_wg);
/* Times every (workgroup shape, persistent grid of 1..8 workgroups per CU)
 * on this batch, keeps the fastest and records it in the tuning file
 * ($METALMOM_AMD_TUNE_FILE or ~/.cache/metalmom_amd_tuning.json); new plans
 * of the same configuration start from the recorded shape and grid. */
int32_t mm_batch_autotune(mm_batch b, const void* d_signal, float* d_out, int32_t iters,
                          int32_t* best_wg_per_cu, float* best_ms);
/* Average device time of the fused phase-1 kernel and of the whole step over
 * `iters` runs, from HIP events on the stream the kernels run on. */
int32_t mm_batch_time(mm_batch b, const void* d_signal, float* d_out, void* hip_stream, int32_t iters,
                      float* kernel_ms, float* step_ms);
int32_t mm_batch_destroy(mm_batch b);

/* Request server with continuous batching. Requests of any kind/config are
 * admitted FIFO into the next device step while the HBM budget (bytes of
 * device input + output per request) and a free stream allow it; up to three
 * steps are in flight, each one persistent launch per configuration group.
 * Results are bitwise identical to the same clip run alone. Thread-safe. */
typedef void* mm_server;
int32_t mm_server_create(mm_context ctx, int64_t hbm_budget_bytes, int32_t max_batch, mm_server* out);
/* Copies the clip and queues it; returns the request id (> 0) or a status. */
int64_t mm_server_submit(mm_server s, const MMSpecParams* params, const float* signal_data,
                         int64_t signal_length);
/* Interleaved little-endian PCM sample formats (integers scaled by
 * 2^-(bits-1), unsigned 8-bit centred at 128). */
#define MM_PCM_U8 1
#define MM_PCM_S16 2
#define MM_PCM_S24 3 /* packed, 3 bytes per sample */
#define MM_PCM_S32 4
#define MM_PCM_F32 5
#define MM_PCM_F64 6
/* Raw-PCM request: `frames` interleaved frames of `channels` (1..64) samples
 * in `pcm_format` at source_sr. On the GPU, in the request's step, each frame
 * becomes the mean of its channels (mono != 0) or its first channel, and the
 * signal is resampled to params->sample_rate with mm_resample's converter when
 * the rates differ; the result is the analysis of that clip, bitwise equal to
 * running it alone. Copies the data and queues it; returns the request id. */
int64_t mm_server_submit_pcm(mm_server s, const MMSpecParams* params, const void* pcm, int64_t frames,
                             int32_t channels, int32_t pcm_format, int32_t source_sr, int32_t mono);
/* 1: done (out filled, latency = completion - submission in seconds, id
 * released), 0: pending, < 0: the request's error status. */
int32_t mm_server_poll(mm_server s, int64_t request_id, MMBuffer* out, double* latency_s);
/* Like mm_server_poll but blocks until the request is done. */
int32_t mm_server_wait(mm_server s, int64_t request_id, MMBuffer* out, double* latency_s);
int32_t mm_server_stats(mm_server s, int64_t* submitted, int64_t* completed, int64_t* steps,
                        int64_t* frames);
/* Device memory the server's streams hold for their steps (inputs, outputs,
 * work lists, scratch; not the shared tables): at most 2.5 x the HBM budget
 * while steps run, at most the budget once they are done. No counterpart in
 * the reference (it has no server). */
int32_t mm_server_device_bytes(mm_server s, int64_t* bytes);
/* Finishes queued work, then releases the server. */
int32_t mm_server_destroy(mm_server s);

/* Device description: name, gfx arch and compute-unit count. */
int32_t mm_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len, int32_t* n_cu);

#ifdef __cplusplus
}
#endif

#endif /* METALMOM_H */
