/*
 * bcsim_c.h -- C-ABI of the MI355X (gfx950) OpenBCSim "gpu" simulation path.
 *
 * Library: libbcsim_hip.so (openbcsim_amd/libbcsim_hip.so after building).
 *
 * Each entry point replaces one member of the reference's plugin interface
 * bcsim::IAlgorithm (src/core/LibBCSim.hpp:42-91) as implemented by
 * GpuAlgorithm (src/core/algorithm/GpuAlgorithm.cpp); the citation on each
 * declaration names the reference member it stands for.  Conventions:
 *
 *   - Every function returns a bcsim_status; BCSIM_OK is 0.  On failure the
 *     calling thread's message is available from bcsim_last_error(); the C++
 *     facade (include/bcsim/LibBCSim.hpp) rethrows it as std::runtime_error,
 *     which is what the reference throws for the same conditions.
 *   - Plain pointers and sizes only.  All inputs are copied before the call
 *     returns; the caller keeps ownership of its buffers.
 *   - A handle is used from one thread at a time.  Each handle owns one HIP
 *     stream on its device; device selection is per handle ("gpu_device").
 *   - Vectors are float triples; arrays of them are [n][3] row-major.
 *   - IQ output is complex64 interleaved (re, im), laid out [line][sample].
 */
#ifndef BCSIM_C_H
#define BCSIM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define BCSIM_API __attribute__((visibility("default")))
#else
#define BCSIM_API
#endif

typedef struct bcsim_sim bcsim_sim; /* opaque simulator handle */
typedef int bcsim_status;

#define BCSIM_OK 0
#define BCSIM_ERR_INVALID_ARGUMENT 1 /* illegal value, shape or key */
#define BCSIM_ERR_INVALID_STATE 2    /* call order or missing configuration */
#define BCSIM_ERR_DEVICE 3           /* HIP runtime or kernel launch failure */
#define BCSIM_ERR_OUT_OF_MEMORY 4    /* device or pinned-host allocation failed */
#define BCSIM_ERR_CAPACITY 5         /* caller's output buffer is too small */
#define BCSIM_ERR_INTERNAL 6         /* logic error inside the library */

/* Message of the last failed call on this thread ("" if none). */
BCSIM_API const char* bcsim_last_error(void);

/* Factory.  Replaces bcsim::Create (src/core/LibBCSim.hpp:97, LibBCSim.cpp:38-48).
 * sim_type "gpu" creates the HIP implementation.  Any other type fails with
 * BCSIM_ERR_INVALID_ARGUMENT ("Illegal algorithm type: ..."); the reference's
 * "cpu" algorithm is not part of this library. */
BCSIM_API bcsim_status bcsim_create(const char* sim_type, bcsim_sim** out);
BCSIM_API bcsim_status bcsim_create_simulator(const char* sim_type, bcsim_sim** out); /* alias */
BCSIM_API bcsim_status bcsim_destroy(bcsim_sim* sim);

/* IAlgorithm::set_parameter (LibBCSim.hpp:50; BaseAlgorithm.cpp:48-87,
 * GpuAlgorithm.cpp:68-107).  Keys: verbose, sound_speed, noise_amplitude,
 * use_arc_projection, radial_decimation, phase_delay, gpu_device,
 * cuda_streams, threads_per_block, store_kernel_details; and keys of this
 * implementation: noise_seed (counter-based noise stream); projection_mode
 * ("tiled", the culled tile kernel, default; "direct", every line-scatterer
 * pair, for checking); autotune (on/off, default off: time candidate launch
 * shapes once per shape class, process-wide, and use the fastest); fft_radix
 * (4 or 8; 0 = tuned or default); tile_lines, tile_block, tile_slices,
 * tile_windows (tile-kernel launch shape; 0 = tuned or heuristic); tile_order
 * (on/off, default on: large projection launches run their blocks heaviest
 * first, by the previous frame's counts); chunk_order (on/off, default on: the
 * projection visits the chunks near to far from the scan origin).  Neither
 * changes a frame beyond float summation order.  weight_cutoff (default 1e-8,
 * in [0, 1)): beam-profile weights below this fraction of the profile's peak
 * magnitude (1 for the Gaussian, the largest |sample| of a LUT) count as zero
 * and their pairs are culled; each dropped pair changes its bin by less than
 * the cutoff times a full-weight echo of the same amplitude; 0 sums every pair
 * whose weight is not exactly zero, as the reference does.  projection_counters
 * (on/off, default off): every projection launch counts its profile-weight
 * evaluations, LDS bin additions and culled-in (chunk, line) pairs into the
 * debug keys projection_weight_evals, projection_lds_adds and
 * projection_chunk_line_pairs (one value per launch).  bin_accumulation (auto,
 * fixed or float; default auto): the projection's LDS bins as 64-bit fixed point
 * (integer LDS atomics; each contribution rounded to a power-of-two grid chosen
 * per launch from the dataset's amplitudes, about 1e-12 of a typical one, then
 * summed exactly) or float; auto tunes the choice where tuning is on and is
 * otherwise fixed for the Gaussian profile and float for a LUT.  Datasets whose
 * amplitudes the grid cannot resolve (a non-finite amplitude, or a dynamic
 * range beyond ~2^39/n) use float bins whatever the setting.
 * Read-only keys of bcsim_get_parameter: launch_plan (shapes of the last frame)
 * and autotune_report (every tuned class with its trial timings). */
BCSIM_API bcsim_status bcsim_set_parameter(bcsim_sim* sim, const char* key, const char* value);

/* IAlgorithm::get_parameter (LibBCSim.hpp:53; GpuAlgorithm.cpp:761-773).
 * Keys: num_cuda_devices, cur_device_name, plus every key accepted by
 * bcsim_set_parameter.  Writes a NUL-terminated string of at most buf_len bytes. */
BCSIM_API bcsim_status bcsim_get_parameter(bcsim_sim* sim, const char* key, char* buf, size_t buf_len);

/* IAlgorithm::set_excitation (LibBCSim.hpp:71; GpuAlgorithm.cpp:379-408).
 * samples[n] float32; 0 <= center_index < n; fs and demod_freq in Hz. */
BCSIM_API bcsim_status bcsim_set_excitation(bcsim_sim* sim, const float* samples, int n,
                                            int center_index, float fs, float demod_freq);

/* IAlgorithm::set_scan_sequence (LibBCSim.hpp:68; GpuAlgorithm.cpp:411-451) with
 * the Scanline/ScanSequence data (ScanSequence.hpp:39-123).  Must follow
 * set_excitation.  Each line is validated like the Scanline constructor
 * (ScanSequence.cpp:37-70: orthonormal to 1e-4); the elevational direction is
 * lateral x direction.  timestamps[num_lines]; line_length in metres. */
BCSIM_API bcsim_status bcsim_set_scan_sequence(bcsim_sim* sim, int num_lines, const float* origins,
                                               const float* directions, const float* lateral_dirs,
                                               const float* timestamps, float line_length,
                                               int all_timestamps_equal);

/* IAlgorithm::clear_fixed_scatterers / add_fixed_scatterers (LibBCSim.hpp:56-59;
 * GpuAlgorithm.cpp:607-614, GpuScatterers.cpp:42-46).  xyza is [n][4] float32
 * (x, y, z, amplitude), the layout of pyrfsim's add_fixed_scatterers
 * (src/python/PythonInterface.cpp:77-94).  set_* = clear_* then add_*. */
BCSIM_API bcsim_status bcsim_clear_fixed_scatterers(bcsim_sim* sim);
BCSIM_API bcsim_status bcsim_add_fixed_scatterers(bcsim_sim* sim, const float* xyza, size_t n);
BCSIM_API bcsim_status bcsim_set_fixed_scatterers(bcsim_sim* sim, const float* xyza, size_t n);

/* IAlgorithm::clear_spline_scatterers / add_spline_scatterers (LibBCSim.hpp:62-65;
 * GpuScatterers.cpp:207-272; SplineScatterers BCSimConfig.hpp:126-165).
 * degree <= 4; knots[n_knots] with n_knots == num_cp + degree + 1;
 * control_points [n][num_cp][3]; amplitudes [n]. */
BCSIM_API bcsim_status bcsim_clear_spline_scatterers(bcsim_sim* sim);
BCSIM_API bcsim_status bcsim_add_spline_scatterers(bcsim_sim* sim, int degree, const float* knots,
                                                   int n_knots, const float* control_points,
                                                   const float* amplitudes, size_t n, int num_cp);
BCSIM_API bcsim_status bcsim_set_spline_scatterers(bcsim_sim* sim, int degree, const float* knots,
                                                   int n_knots, const float* control_points,
                                                   const float* amplitudes, size_t n, int num_cp);

/* IAlgorithm::set_analytical_profile with a GaussianBeamProfile
 * (LibBCSim.hpp:74; GpuAlgorithm.cpp:453-461; BeamProfile.cpp:34-56). */
BCSIM_API bcsim_status bcsim_set_analytical_profile(bcsim_sim* sim, float sigma_lateral,
                                                    float sigma_elevational);

/* IAlgorithm::set_lookup_profile with a LUTBeamProfile (LibBCSim.hpp:77;
 * GpuAlgorithm.cpp:463-521; BeamProfile.cpp:58-137).  samples is
 * [n_rad][n_lat][n_ele] float32 (index e + n_ele*l + n_lat*n_ele*r,
 * BeamProfile.hpp:127-129); every n >= 2.  Sampled with the reference CPU's
 * trilinear rule, zero outside the table. */
BCSIM_API bcsim_status bcsim_set_lut_profile(bcsim_sim* sim, float r_min, float r_max, float l_min,
                                             float l_max, float e_min, float e_max,
                                             const float* samples, int n_rad, int n_lat, int n_ele);

/* Shape of the next simulate_lines output (no reference counterpart; the
 * reference returns std::vector<std::vector<>>). */
BCSIM_API bcsim_status bcsim_get_output_shape(bcsim_sim* sim, int* num_lines, int* num_samples);

/* IAlgorithm::simulate_lines (LibBCSim.hpp:81; GpuAlgorithm.cpp:145-377).
 * Writes num_lines*num_samples complex64 values into out (capacity counted in
 * complex elements) and returns when they are there. */
BCSIM_API bcsim_status bcsim_simulate_lines(bcsim_sim* sim, float* out, size_t capacity,
                                            int* num_lines, int* num_samples);

/* Device-resident variant of simulate_lines: enqueues the frame on the
 * handle's stream with the IQ written to device memory d_out (complex64,
 * [line][sample]) and returns without waiting.  bcsim_synchronize waits. */
BCSIM_API bcsim_status bcsim_simulate_lines_device(bcsim_sim* sim, void* d_out, size_t capacity,
                                                   int* num_lines, int* num_samples);
BCSIM_API bcsim_status bcsim_synchronize(bcsim_sim* sim);

/* Allocation-free variant of bcsim_simulate_lines_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a frame
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or other allocating routine: the
 * frame's projection bins (and the rendered points of all-timestamps-equal
 * spline sets) live in a fixed scratch array of the library, 32 MiB per GPU, held
 * until the process ends and shared by every call on that GPU, which keeps the
 * last call's contents.  Calls that overlap on one GPU (concurrent calls from
 * several handles or threads, or frames still in flight on other streams) race on
 * it: serialise them.  Frames needing more scratch than 32 MiB (L * num_rf * 4
 * bytes of bins, 8 with phase delay, plus 16 bytes per rendered spline) fail with
 * BCSIM_ERR_CAPACITY.  The handle's tables must already be built for its current
 * configuration: after configuring it (or changing its configuration) run one
 * frame through bcsim_simulate_lines_device or bcsim_simulate_lines, else
 * BCSIM_ERR_INVALID_STATE.  Launch shapes are the user-set or heuristic ones (no
 * tuning trials), blocks run in index order, and store_kernel_details and
 * projection_counters record nothing; the IQ equals that of
 * bcsim_simulate_lines_device up to float summation order.
 * bcsim_debug_get_projection fails after such a frame. */
BCSIM_API bcsim_status bcsim_simulate_lines_device_noalloc(bcsim_sim* sim, void* d_out, size_t capacity,
                                                           int* num_lines, int* num_samples);
/* The handle's hipStream_t, as void*. */
BCSIM_API bcsim_status bcsim_get_stream(bcsim_sim* sim, void** stream);

/* IAlgorithm::get_debug_data (LibBCSim.hpp:84; BaseAlgorithm.cpp:89-94).
 * With store_kernel_details on, keys hold per-launch milliseconds measured with
 * HIP events on the handle's stream: kernel_memset_ms (0 when no fill ran),
 * fixed_projection_kernel_ms, spline_projection_kernel_ms, kernel_render_ms,
 * kernel_convolve_demodulate_ms, kernel_memcpy_ms (0 for device output); and,
 * as the reference records them (GpuAlgorithm.cpp:193-358), stream_numbers (0
 * for every line: one stream per frame) and kernel_forward_fft_ms (the fused
 * convolution kernel's time) with kernel_multiply_fft_ms, kernel_inverse_fft_ms
 * and kernel_demodulate_ms 0 (fused into that kernel).  *count receives the number of values; at most capacity are
 * written.  Unknown keys fail like the reference ("invalid debug data key"). */
BCSIM_API bcsim_status bcsim_get_debug_data(bcsim_sim* sim, const char* key, double* out,
                                            size_t capacity, size_t* count);

/* IAlgorithm::get_total_num_scatterers (LibBCSim.hpp:87; GpuAlgorithm.cpp:755-759). */
BCSIM_API bcsim_status bcsim_get_total_num_scatterers(bcsim_sim* sim, size_t* out);

/* IAlgorithm::set_logger (LibBCSim.hpp:90; ILog BCSimConfig.hpp:57-69).
 * level is ILog::LogType (0 DEBUG, 1 INFO, 2 WARNING, 3 FATAL); NULL removes it. */
typedef void (*bcsim_log_fn)(int level, const char* msg, void* user);
BCSIM_API bcsim_status bcsim_set_logger(bcsim_sim* sim, bcsim_log_fn fn, void* user);

/* Time-projection bins of the last simulated frame, [line][num_rf] complex64.
 * No reference counterpart (the reference keeps them internal,
 * GpuAlgorithm.cpp:205-265); exported so the projection stage can be checked
 * on its own.  With out == NULL only the shape is returned. */
BCSIM_API bcsim_status bcsim_debug_get_projection(bcsim_sim* sim, float* out, size_t capacity,
                                                  int* num_lines, int* num_rf);

/* ------------------------------------------------------------------------------------------
 * Batched frames.  No reference counterpart: the reference simulates one frame per
 * simulate_lines call with per-line launches (GpuAlgorithm.cpp:189-360).  A batch
 * holds configured simulator handles (jobs); bcsim_batch_simulate* runs one frame of
 * every job with one persistent projection launch and one persistent convolution
 * launch, captured into a HIP graph and replayed while no job changes.  Each job's
 * IQ is written at its offset of one complex64 output (bcsim_batch_get_layout).
 * A handle in a batch must not run frames of its own concurrently with the batch.
 */
typedef struct bcsim_batch bcsim_batch;

BCSIM_API bcsim_status bcsim_batch_create(bcsim_batch** out);
BCSIM_API bcsim_status bcsim_batch_destroy(bcsim_batch* batch);
/* Keys: use_graph (on/off, default on), store_kernel_details (on/off: records
 * batch_frame_ms with HIP events, waiting for each frame), autotune (on/off,
 * default on: each job's shape class is tuned before its first batched frame),
 * tile_lines (lines per tile for every job, 0 = each job's plan),
 * items_per_workgroup (projection work items per persistent workgroup, default 4). */
BCSIM_API bcsim_status bcsim_batch_set_parameter(bcsim_batch* batch, const char* key, const char* value);
BCSIM_API bcsim_status bcsim_batch_add(bcsim_batch* batch, bcsim_sim* sim, int* job_index);
BCSIM_API bcsim_status bcsim_batch_clear(bcsim_batch* batch);
/* total complex elements of the output; per job (up to capacity jobs) its offset,
 * lines and samples per line. */
BCSIM_API bcsim_status bcsim_batch_get_layout(bcsim_batch* batch, size_t* total, int* num_jobs,
                                              size_t* offsets, int* lines, int* samples, int capacity);
/* Enqueue one frame of every job, IQ to device memory d_out; returns without waiting. */
BCSIM_API bcsim_status bcsim_batch_simulate_device(bcsim_batch* batch, void* d_out, size_t capacity);
/* One frame of every job into host memory; returns when the data is there. */
BCSIM_API bcsim_status bcsim_batch_simulate(bcsim_batch* batch, float* out, size_t capacity);
BCSIM_API bcsim_status bcsim_batch_synchronize(bcsim_batch* batch);
BCSIM_API bcsim_status bcsim_batch_get_debug_data(bcsim_batch* batch, const char* key, double* out,
                                                  size_t capacity, size_t* count);

/* ------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, whose frames are one
 * simulate_lines call each, GpuAlgorithm.cpp:189-360; its GUI packet loop,
 * MainWindow.cpp:666-695, is the usage pattern served here).
 *
 * A server keeps phantoms resident in device memory and takes frame requests
 * (phantom, scan geometry, time t) from any thread.  A scheduler thread admits
 * requests in arrival order while their frame buffers fit the memory budget and
 * launches every admitted request as one batched frame; the next batch is formed
 * and launched while the previous one runs (frame_contexts frames in flight).
 * A request of more than piece_lines lines runs as pieces (consecutive line
 * ranges) in successive frames, so that its first lines cross PCIe while the
 * rest compute.  A request's IQ equals what a simulator of that phantom with
 * that scan (all timestamps t) returns from simulate_lines, up to float
 * summation order.
 */
typedef struct bcsim_server bcsim_server;

BCSIM_API bcsim_status bcsim_server_create(bcsim_server** out);
/* Fails requests that have not run; waits for frames in flight. */
BCSIM_API bcsim_status bcsim_server_destroy(bcsim_server* server);
/* Keys: memory_budget_bytes (device bytes for phantoms, frame buffers and
 * outputs; default 90% of free device memory when the first phantom is added),
 * max_jobs_per_frame (default 64), autotune (on/off, default on),
 * piece_lines (default 1024; 0: never split a request), frame_contexts
 * (frames in flight, 1-4, default 2), reset_stats (any value: clears the
 * statistics). */
BCSIM_API bcsim_status bcsim_server_set_parameter(bcsim_server* server, const char* key, const char* value);
/* Snapshot of a configured simulator (excitation, parameters, profile,
 * scatterer datasets): its datasets stay resident, shared with the caller's
 * simulator, which may be changed or destroyed afterwards. */
BCSIM_API bcsim_status bcsim_server_add_phantom(bcsim_server* server, bcsim_sim* phantom, int* phantom_id);
/* A scan geometry: num_lines origins, directions and lateral directions
 * (float[3] each, as bcsim_set_scan_sequence) and the line length in metres. */
BCSIM_API bcsim_status bcsim_server_add_geometry(bcsim_server* server, int num_lines, const float* origins,
                                                 const float* directions, const float* lateral_dirs,
                                                 float line_length, int* geometry_id);
BCSIM_API bcsim_status bcsim_server_get_output_shape(bcsim_server* server, int phantom_id, int geometry_id,
                                                     int* num_lines, int* num_samples);
/* Queue one frame; out (complex64 [lines][samples], capacity complex elements)
 * must stay valid until bcsim_server_wait reports the request.  When out lies in
 * a buffer of bcsim_server_alloc_host, the IQ is copied there by DMA and wait
 * copies nothing. */
BCSIM_API bcsim_status bcsim_server_submit(bcsim_server* server, int phantom_id, int geometry_id, float t,
                                           float* out, size_t capacity, uint64_t* request_id);
/* Wait (timeout_ms < 0: without limit) for a request; returns its status and
 * releases its record.  The IQ, staged in pinned memory when the frame ends, is
 * copied into the caller's buffer here, in the calling thread.  latency_ms is
 * the time from submit to the IQ in the caller's buffer for a caller that
 * waits from submit on (staging time plus this copy). */
BCSIM_API bcsim_status bcsim_server_wait(bcsim_server* server, uint64_t request_id, int timeout_ms,
                                         int* num_lines, int* num_samples, double* latency_ms);
/* Keys: latency_ms, queue_ms (submit to launch), copy_ms (the copy in wait)
 * per waited request;
 * frame_jobs, frame_lines, launch_ms (host time to prepare and enqueue) per
 * launched frame; lines_completed, busy_span_ms
 * (first submit to last completion), budget_bytes, used_bytes, resident_bytes,
 * idle_slots, pending (one value each).  out NULL: count only. */
BCSIM_API bcsim_status bcsim_server_get_stats(bcsim_server* server, const char* key, double* out,
                                              size_t capacity, size_t* count);
/* Pinned host memory for request outputs (no counterpart in the reference): the
 * frame's IQ goes straight into it by DMA, without the staging copy.  Freed by
 * bcsim_server_free_host (an error while an unreported request writes into it)
 * or with the server. */
BCSIM_API bcsim_status bcsim_server_alloc_host(bcsim_server* server, size_t bytes, void** ptr);
BCSIM_API bcsim_status bcsim_server_free_host(bcsim_server* server, void* ptr);

#ifdef __cplusplus
}
#endif
#endif /* BCSIM_C_H */
