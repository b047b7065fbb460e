/*
 * cpsz — critical-point-preserving lossy compression of 2D/3D float vector
 * fields on AMD Instinct MI355X (gfx950). C ABI of libcpsz_hip.so.
 *
 * Streams are byte-compatible with the reference CPU library (bzhan166/cpSZ):
 * a stream written here decodes with the reference decoder and vice versa.
 * ZSTD stays a separate host-side stage, as in the reference drivers.
 *
 * All entry points return 0 on success or a negative CPSZ_ERR_* code; the
 * message of the calling thread's last failure is cpsz_last_error(). Calls are
 * thread-safe: each host thread gets its own per-device context (HIP stream and
 * scratch buffers).
 */
#ifndef CPSZ_H_
#define CPSZ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CPSZ_OK = 0,
  CPSZ_ERR_ARG = -1,
  CPSZ_ERR_HIP = -2,
  CPSZ_ERR_ALLOC = -3,
  CPSZ_ERR_STREAM = -4,
  CPSZ_ERR_UNSUPPORTED = -5,
  CPSZ_ERR_NONFINITE = -6,
  CPSZ_ERR_INTERNAL = -7
};

/* Message of the calling thread's last failed call ("" if none). */
const char* cpsz_last_error(void);
/* Library version string. */
const char* cpsz_version(void);
/* Number of visible HIP devices (0 when none). */
int cpsz_device_count(void);
/* Releases a buffer returned by this library. */
void cpsz_free(void* p);

/* ---- host-buffer codec (what the drop-in C++ templates call) ---------- */

/* Replaces sz_compress_cp_preserve_2d_offline<float>
 * (reference include/sz_compress_cp_preserve_2d.hpp:8-10; implementation
 * src/sz_compress_cp_preserve_2d.cpp:175-274). U, V: r1 x r2 row-major host
 * arrays. *out is malloc'd (release with free() or cpsz_free). */
int cpsz_compress_2d_f32(const float* U, const float* V, size_t r1, size_t r2, double max_pwr_eb,
                         unsigned char** out, size_t* out_size);

/* Replaces sz_decompress_cp_preserve_2d_offline<float>
 * (reference include/sz_decompress_cp_preserve_2d.hpp:8-11; implementation
 * src/sz_decompress_cp_preserve_2d.cpp:8-44). U, V: caller buffers of r1*r2. */
int cpsz_decompress_2d_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2,
                           float* U, float* V);

/* Replace sz_compress_cp_preserve_2d_offline<double> and
 * sz_decompress_cp_preserve_2d_offline<double> (reference explicit
 * instantiations src/sz_compress_cp_preserve_2d.cpp:280-282,
 * src/sz_decompress_cp_preserve_2d.cpp:49-52): the double arithmetic of the
 * reference's double code, 8-byte unpredictable values in the stream. */
int cpsz_compress_2d_f64(const double* U, const double* V, size_t r1, size_t r2, double max_pwr_eb,
                         unsigned char** out, size_t* out_size);
int cpsz_decompress_2d_f64(const unsigned char* in, size_t in_size, size_t r1, size_t r2, double* U,
                           double* V);
int cpsz_compress_2d_f64_dev(int device, void* stream, const double* dU, const double* dV, size_t r1,
                             size_t r2, double max_pwr_eb, const unsigned char** d_out, size_t* out_size);
int cpsz_decompress_2d_f64_dev(int device, void* stream, const unsigned char* d_in, size_t in_size,
                               size_t r1, size_t r2, double* dU, double* dV);

/* Replaces sz_compress_cp_preserve_2d_offline_log<float>
 * (reference include/sz_compress_cp_preserve_2d.hpp:12-14; implementation
 * src/sz_compress_cp_preserve_2d.cpp:285-374): the offline bounds in the log
 * domain, the codec on log2f|x|, sign bitmaps in the stream. */
int cpsz_compress_2d_log_f32(const float* U, const float* V, size_t r1, size_t r2, double max_pwr_eb,
                             unsigned char** out, size_t* out_size);
/* Replaces sz_decompress_cp_preserve_2d_offline_log<float>
 * (reference include/sz_decompress_cp_preserve_2d.hpp:12-14; implementation
 * src/sz_decompress_cp_preserve_2d.cpp:54-99). */
int cpsz_decompress_2d_log_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2, float* U,
                               float* V);
int cpsz_compress_2d_log_f32_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                 size_t r2, double max_pwr_eb, const unsigned char** d_out, size_t* out_size);
int cpsz_decompress_2d_log_f32_dev(int device, void* stream, const unsigned char* d_in, size_t in_size,
                                   size_t r1, size_t r2, float* dU, float* dV);
/* Total byte size of a 2D offline_log stream from its own framing. */
int cpsz_stream_size_2d_log(const unsigned char* in, size_t r1, size_t r2, size_t* size);

/* Replaces sz_compress_cp_preserve_3d_offline_log<float>
 * (reference include/sz_compress_cp_preserve_3d.hpp:6-8; implementation
 * src/sz_compress_cp_preserve_3d.cpp:360-581). r1 slowest .. r3 fastest. */
int cpsz_compress_3d_f32(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                         size_t r3, double max_pwr_eb, unsigned char** out, size_t* out_size);

/* Replaces sz_decompress_cp_preserve_3d_online_log<float>
 * (reference include/sz_decompress_cp_preserve_3d.hpp:8-11; implementation
 * src/sz_decompress_cp_preserve_3d.cpp:7-110). */
int cpsz_decompress_3d_f32(const unsigned char* in, size_t in_size, size_t r1, size_t r2,
                           size_t r3, float* U, float* V, float* W);

/* Total byte size of a stream from its own framing (the reference decoders
 * take no size argument; these read the section headers to find the end). */
int cpsz_stream_size_2d(const unsigned char* in, size_t* size);
int cpsz_stream_size_3d(const unsigned char* in, size_t r1, size_t r2, size_t r3, size_t* size);

/* ---- device-resident codec ------------------------------------------------
 * Inputs and outputs live in HBM of `device`; `stream` is a hipStream_t (NULL:
 * the context's own stream). The compressed stream is returned in a
 * library-owned device buffer valid until the next compress call of this
 * thread on this device. */
int cpsz_compress_2d_f32_dev(int device, void* stream, const float* dU, const float* dV,
                             size_t r1, size_t r2, double max_pwr_eb,
                             const unsigned char** d_out, size_t* out_size);
int cpsz_decompress_2d_f32_dev(int device, void* stream, const unsigned char* d_in,
                               size_t in_size, size_t r1, size_t r2, float* dU, float* dV);
int cpsz_compress_3d_f32_dev(int device, void* stream, const float* dU, const float* dV,
                             const float* dW, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                             const unsigned char** d_out, size_t* out_size);
int cpsz_decompress_3d_f32_dev(int device, void* stream, const unsigned char* d_in,
                               size_t in_size, size_t r1, size_t r2, size_t r3, float* dU,
                               float* dV, float* dW);

/* ---- host lossless stage ---------------------------------------------------
 * Replaces sz_lossless_compress / sz_lossless_decompress with ZSTD
 * (reference src/sz_lossless.cpp:11-45; level 3 in the drivers). */
int cpsz_lossless_compress(int level, const unsigned char* in, size_t n, unsigned char** out,
                           size_t* out_size);
int cpsz_lossless_decompress(const unsigned char* in, size_t n, unsigned char* out,
                             size_t out_size);
/* Replaces is_lossless_compressed_data (reference src/sz_lossless.cpp:3-9):
 * 1 when `in` starts with a ZSTD frame whose header parses, else 0. */
int cpsz_lossless_is_frame(const unsigned char* in, size_t n);
/* sz_lossless_compress (src/sz_lossless.cpp:11-29) of a stream that is in
 * device memory (as the _dev compressors return it): the bytes travel to
 * pinned host memory in 8 MiB chunks on `stream` while the host compresses
 * the chunks that have arrived (zstd's block-level API over the contiguous
 * buffer, whole 128 KiB blocks per chunk); the frame is byte-identical to the
 * reference's one-shot call at `level`.
 * *out is malloc'd (free with cpsz_free). */
int cpsz_lossless_compress_dev(int device, void* stream, int level, const unsigned char* d_in, size_t n,
                               unsigned char** out, size_t* out_size);
/* The same for a batch of device streams d_base + off[k] (size[k] bytes, as
 * the batched compressors return them): one frame per field, each identical
 * to the reference's sz_lossless_compress of that field's stream, compressed
 * by up to 16 host threads as the fields' bytes arrive. */
int cpsz_lossless_compress_batch_dev(int device, void* stream, int level, const unsigned char* d_base,
                                     const size_t* off, const size_t* size, size_t nf, unsigned char** outs,
                                     size_t* out_sizes);

/* Replaces the decoding in sz_lossless_decompress65536bytes (reference
 * src/sz_lossless.cpp:47-62): decodes the first out_size bytes of the frame
 * (fewer if the frame is shorter) into `out`, zero-filling the rest of it;
 * *produced = bytes decoded. The reference reads a stream header this way. */
int cpsz_lossless_decompress_prefix(const unsigned char* in, size_t n, unsigned char* out,
                                    size_t out_size, size_t* produced);

/* ---- critical points (0-CP-error judge) ----------------------------------
 * Critical points per simplex, decided with exact sign predicates on the GPU.
 * 2D: interior cells (reference test/cp_extraction_2d.cpp:126-145), ids
 * 2*(i*(r2-1)+j)(+1); 3D: all cells (test/cp_extraction_3d.cpp:174-206), ids
 * 6*(vertex index)+tet. Writes up to `cap` (id, type) pairs sorted by id and
 * the total count to *count. */
int cpsz_critical_points_2d_dev(int device, void* stream, const float* dU, const float* dV,
                                size_t r1, size_t r2, int64_t* ids, int32_t* types, size_t cap,
                                size_t* count);
/* The 2D critical-point judge without allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same critical points as cpsz_critical_points_2d_dev, but it calls no
 * hipMalloc/hipFree and creates no context: its scratch is a fixed device
 * array of 12 MiB + 256 B per GPU (held until the process ends), shared by
 * every call on that GPU, so calls that overlap on one GPU race. `stream` may
 * be NULL (the null stream). Limits: at most 2^20 critical points (else
 * CPSZ_ERR_UNSUPPORTED), r1, r2 < 2^28; ids/types are host arrays of `cap`
 * entries and cap must be >= the count (else CPSZ_ERR_ARG, *count set). */
int cpsz_critical_points_2d_noalloc_dev(int device, void* stream, const float* dU, const float* dV,
                                        size_t r1, size_t r2, int64_t* ids, int32_t* types, size_t cap,
                                        size_t* count);
int cpsz_critical_points_3d_dev(int device, void* stream, const float* dU, const float* dV,
                                const float* dW, size_t r1, size_t r2, size_t r3, int64_t* ids,
                                int32_t* types, size_t cap, size_t* count);

/* The two judges above under Simulation of Simplicity (Edelsbrunner & Muecke
 * 1990): vertex vectors are perturbed symbolically by their global index, so
 * a zero of the field lies in exactly one simplex, never on a face, and zero
 * vertex vectors take part like any other. Same simplices, ids and type codes
 * as the closed-simplex judges, which follow the reference's extractor
 * (test/cp_extraction_2d.cpp:62-147, test/cp_extraction_3d.cpp:79-209); the
 * perturbation has no counterpart in the reference (the north-star's exact
 * sign/SoS cell test). */
int cpsz_critical_points_2d_sos_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                    size_t r2, int64_t* ids, int32_t* types, size_t cap, size_t* count);
int cpsz_critical_points_3d_sos_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                                    size_t r1, size_t r2, size_t r3, int64_t* ids, int32_t* types, size_t cap,
                                    size_t* count);
/* Sign (+-1, never 0) of the d x d determinant (d = 2 or 3) of the rows
 * rows[0..d*d) of vertices with distinct global indices idx[0..d) under the
 * same perturbation (host evaluation of the device predicate, for tests). */
int cpsz_sos_det(int d, const float* rows, const int64_t* idx, int* sign);

/* ---- host-side diagnostics (no GPU needed) -------------------------------- */
/* Copies the eb-id threshold tables the device kernels use: q2d[k] is the
 * smallest |U|*eb with 2D id >= k, q3d[k] the smallest vertex bound with 3D
 * id >= k (both from this host's libm, as the reference computes ids). Returns
 * the table lengths through n2d / n3d. */
int cpsz_eb_tables(double* q2d, int cap2d, int* n2d, double* q3d, int cap3d, int* n3d);
/* The 2D offline_log eb-code thresholds: q[k] is the smallest vertex bound e
 * whose code (0 = zero precision, c >= 1 = id c - 1 of
 * eb_exponential_quantize(log2(1 + e), 2, 1, DBL_EPSILON)) is >= k. */
int cpsz_eb_table_2d_log(double* q, int cap, int* n);
/* Builds the Huffman tree the encoder would emit for the given symbol
 * frequencies and writes the serialized tree bytes and node count. */
int cpsz_huffman_tree(const uint64_t* freq, size_t nbins, uint32_t state_num, uint8_t* tree,
                      size_t cap, size_t* tree_size, size_t* node_count);
/* The same tree built on the GPU (one wave, the reference heap's discipline)
 * on `device`; codes/lens (nbins entries each, or NULL) receive the dense
 * codebook the encoder uses (code left-aligned in 64 bits, length). */
int cpsz_huffman_tree_dev(int device, const uint64_t* freq, size_t nbins, uint32_t state_num, uint8_t* tree,
                          size_t cap, size_t* tree_size, size_t* node_count, uint64_t* codes, uint8_t* lens);
/* Decodes n symbols of an MSB-first Huffman bitstream on the host with the
 * two-level tables the GPU decoder builds from a serialized tree (a check of
 * the table construction). */
int cpsz_huffman_decode_host(const uint8_t* tree, size_t tree_size, size_t node_count, const uint8_t* bits,
                             size_t nbytes, size_t n, uint32_t* out);

/* Builds the host-libm exception tables that let the device reproduce the
 * host's log2f/exp2f bit for bit, and writes them to `path` (NULL: the cache
 * file next to the library). The library loads that cache on first use and
 * rebuilds it when it does not match the running host libm. */
int cpsz_prepare_libm_tables(const char* path);
/* The device's log2f / exp2f evaluation run on the host (for testing):
 * log2f for finite x > 0, exp2f for y in [-99, 128). */
int cpsz_eval_log2f(const float* x, size_t n, float* out);
int cpsz_eval_exp2f(const float* y, size_t n, float* out);

/* Host emulation of the device 2D Lorenzo arithmetic (row-major order):
 * eb codes c (0 = zero precision, else reference id + 1) -> quant codes q and
 * the reconstruction. `wide` selects the double-precision path used for
 * ids above 81. */
int cpsz_lorenzo2d_emulate(const float* x, const uint16_t* c, size_t r1, size_t r2, int wide,
                           uint16_t* q, float* recon);

/* Host emulations of the device 3D arithmetic: per-vertex bounds (the min
 * over incident tetrahedra of the reference's per-tet bound), and the Lorenzo
 * pass from log values L[3][n] and eb codes (255 = non-positive bound, else
 * the id) to per-vertex stream ids and per-component codes q[3][n]. */
int cpsz_eb3d_emulate(const float* U, const float* V, const float* W, size_t r1, size_t r2,
                      size_t r3, double max_pwr_eb, double* eb);
int cpsz_lorenzo3d_emulate(const float* L, const uint8_t* code, size_t r1, size_t r2, size_t r3,
                           uint8_t* vid, uint16_t* q);
/* eb code of each bound as the device derives it (255 or the 3D id). */
int cpsz_eb_code_3d(const double* eb, size_t n, uint8_t* code);

/* ---- batched 2D codec ----------------------------------------------------
 * Many fields of uneven sizes (no counterpart in the reference, whose drivers
 * compress one field per call). Each field's stream is byte-identical to
 * cpsz_compress_2d_f32_dev's. One persistent wavefront launch covers every
 * field; the entropy stages of different fields run concurrently on worker
 * streams. tune != 0 first times the candidate configurations on this batch
 * (outputs do not depend on them) and caches the fastest for batches of the
 * same signature. Streams are returned in one library-owned device buffer
 * (valid until the next batch call of this thread) at offsets[k], sizes[k]. */
int cpsz_compress_2d_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                               const float* const* dV, const size_t* r1, const size_t* r2,
                               double max_pwr_eb, int tune, const unsigned char** d_base,
                               size_t* offsets, size_t* sizes);
int cpsz_decompress_2d_batch_dev(int device, void* stream, size_t nfields,
                                 const unsigned char* const* d_in, const size_t* in_size,
                                 const size_t* r1, const size_t* r2, float* const* dU,
                                 float* const* dV);
/* Configuration the batch calls use for this batch signature: host worker
   threads, persistent-grid workgroups and (device_trees may be null) whether
   the compressor builds its Huffman trees on the GPU (1) or on the host (0). */
int cpsz_batch_2d_config(int device, size_t nfields, const size_t* r1, const size_t* r2,
                         int* workers, int* grid_cap, int* device_trees);
/* Pin that configuration for batches of these shapes on this device (for
   callers that know their best settings, or want the trees on the GPU). */
int cpsz_batch_2d_set_config(int device, size_t nfields, const size_t* r1, const size_t* r2,
                             int workers, int grid_cap, int device_trees);

/* ---- batched 3D codec ----------------------------------------------------
 * Many 3D fields (timesteps, ensemble members; sizes may differ) per call, no
 * counterpart in the reference. Each field's stream is byte-identical to
 * cpsz_compress_3d_f32_dev's. One persistent wavefront launch covers every
 * field; the entropy stages run concurrently on worker streams. Outputs and
 * tuning as for the batched 2D codec. */
int cpsz_compress_3d_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                               const float* const* dV, const float* const* dW, const size_t* r1,
                               const size_t* r2, const size_t* r3, double max_pwr_eb, int tune,
                               const unsigned char** d_base, size_t* offsets, size_t* sizes);
int cpsz_decompress_3d_batch_dev(int device, void* stream, size_t nfields,
                                 const unsigned char* const* d_in, const size_t* in_size,
                                 const size_t* r1, const size_t* r2, const size_t* r3, float* const* dU,
                                 float* const* dV, float* const* dW);
int cpsz_batch_3d_config(int device, size_t nfields, const size_t* r1, const size_t* r2, const size_t* r3,
                         int* workers, int* grid_cap);

/* ---- learned error-bound proposer (2D) -----------------------------------
 * No counterpart in the reference (its closest precedent is the
 * verify-and-halve loop of sz_compress_cp_preserve_2d_online_abs_relax_FN,
 * src/sz_compress_cp_preserve_2d.cpp:1439-1502). A 51-128-128-1 MLP (BF16 on
 * MFMA, FP32 accumulation) proposes per-vertex bound gains; every proposal is
 * verified on the decompressed field by the exact critical-point judge and
 * falls back to the analytic bound where it fails. Streams keep the reference
 * layout. Parameters are a host float array of cpsz_learn_param_count():
 * W1[64][128], b1[128], W2[128][128], b2[128], w3[128], b3. */
size_t cpsz_learn_param_count(void);
int cpsz_learn_init_params(uint64_t seed, float* params);
/* The plain method's per-vertex bound before quantization (dE: r1*r2 doubles). */
int cpsz_learn_analytic_eb_2d_dev(int device, void* stream, const float* dU, const float* dV, size_t r1,
                                  size_t r2, double max_pwr_eb, double* dE);
/* BF16 features (64 per vertex, 51 used) of vertices d_idx[0..m) (all, row
 * major, when d_idx is NULL). */
int cpsz_learn_features_2d_dev(int device, void* stream, const float* dU, const float* dV,
                               const double* dE, size_t r1, size_t r2, double max_pwr_eb,
                               const uint32_t* d_idx, size_t m, uint16_t* dX);
/* Bisection labels: the largest gain (0..24) that keeps every incident
 * triangle's critical point when the vertex alone moves to any corner of its
 * bound box. */
int cpsz_learn_labels_2d_dev(int device, void* stream, const float* dU, const float* dV,
                             const double* dE, size_t r1, size_t r2, double max_pwr_eb,
                             const uint32_t* d_idx, size_t m, uint8_t* dL);
/* C[M][N] = sum_k A(m,k) B(k,n), BF16 operands with element strides, FP32 C
 * (row-major, ldc); split > 1 accumulates K slices atomically into C. */
int cpsz_gemm_bf16_dev(int device, void* stream, int M, int N, int K, const uint16_t* A, int64_t sam,
                       int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn, float* C, int ldc,
                       int split);
/* Network output for m feature rows; gradient and mean squared error of a
 * batch against targets dT (params, grad and loss in host memory). */
int cpsz_mlp_forward_dev(int device, void* stream, const float* params, const uint16_t* dX, size_t m,
                         float* dY);
int cpsz_mlp_grad_dev(int device, void* stream, const float* params, const uint16_t* dX,
                      const float* dT, size_t m, float* grad, float* loss);
/* Trains on samples_per_field random vertices of each field (features and
 * bisection labels built on the GPU) with Adam; params is the start point and
 * receives the result; loss_per_epoch (may be NULL) gets each epoch's last
 * batch loss. */
int cpsz_learn_train_2d_dev(int device, void* stream, size_t nfields, const float* const* dU,
                            const float* const* dV, const size_t* r1, const size_t* r2,
                            double max_pwr_eb, size_t samples_per_field, int epochs, int batch,
                            float lr, uint64_t seed, float* params, float* loss_per_epoch);
/* Learned-bound compression; the stream is decodable by
 * cpsz_decompress_2d_f32 and the reference decoder, and the decoded field has
 * the original's critical points. iterations: verification rounds; boosted:
 * vertices given a gain; fallback: boosted vertices sent back to the
 * analytic bound. */
int cpsz_compress_2d_learned_f32_dev(int device, void* stream, const float* dU, const float* dV,
                                     size_t r1, size_t r2, double max_pwr_eb, const float* params,
                                     const unsigned char** d_out, size_t* out_size, int* iterations,
                                     size_t* boosted, size_t* fallback);
int cpsz_compress_2d_learned_f32(const float* U, const float* V, size_t r1, size_t r2,
                                 double max_pwr_eb, const float* params, unsigned char** out,
                                 size_t* out_size);

/* ---- learned error-bound proposer (3D) -----------------------------------
 * Same design as the 2D proposer, no counterpart in the reference: an
 * 82-128-128-1 MLP over the 3x3x3 patch of (u, v, w) (divided by the patch's
 * largest magnitude) plus log2(analytic / max_pwr_eb) / 16, BF16 on MFMA with
 * FP32 accumulation. Labels: the largest gain (0..24) that keeps every
 * incident tetrahedron's critical point when the vertex alone moves to any
 * corner of its box x * (1 + eb)^(+-1). Parameters: W1[96][128] (rows 82..95
 * zero), b1, W2[128][128], b2, w3[128], b3. */
size_t cpsz_learn_param_count_3d(void);
int cpsz_learn_init_params_3d(uint64_t seed, float* params);
int cpsz_learn_analytic_eb_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                                  size_t r1, size_t r2, size_t r3, double max_pwr_eb, double* dE);
/* 96 BF16 features per vertex (82 used) of vertices d_idx[0..m) (all when d_idx is NULL). */
int cpsz_learn_features_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                               const double* dE, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                               const uint32_t* d_idx, size_t m, uint16_t* dX);
int cpsz_learn_labels_3d_dev(int device, void* stream, const float* dU, const float* dV, const float* dW,
                             const double* dE, size_t r1, size_t r2, size_t r3, double max_pwr_eb,
                             const uint32_t* d_idx, size_t m, uint8_t* dL);
int cpsz_mlp_forward_3d_dev(int device, void* stream, const float* params, const uint16_t* dX, size_t m,
                            float* dY);
int cpsz_mlp_grad_3d_dev(int device, void* stream, const float* params, const uint16_t* dX, const float* dT,
                         size_t m, float* grad, float* loss);
int cpsz_learn_train_3d_dev(int device, void* stream, size_t nfields, const float* const* dU,
                            const float* const* dV, const float* const* dW, const size_t* r1, const size_t* r2,
                            const size_t* r3, double max_pwr_eb, size_t samples_per_field, int epochs, int batch,
                            float lr, uint64_t seed, float* params, float* loss_per_epoch);
/* Proposal gains of vertices v0 .. v0+m-1 from one fused kernel (features,
 * both hidden layers on BF16 MFMA with FP32 accumulation, output): gain =
 * clamp(floor(net output), 0, 24) into d_gain[0..m). dim 2: the 51-feature
 * net on (dU, dV); dim 3: the 82-feature net on (dU, dV, dW). dE is the
 * analytic bound per vertex (cpsz_learn_analytic_eb_*_dev). No reference
 * counterpart. */
int cpsz_learn_infer_gains_dev(int device, void* stream, int dim, const float* dU, const float* dV,
                               const float* dW, const double* dE, size_t r1, size_t r2, size_t r3,
                               double max_pwr_eb, const float* params, size_t v0, size_t m, uint8_t* d_gain);
/* Learned-bound compression of a batch of 2D fields (no reference
 * counterpart): the single-field learned compressor over many fields, each
 * verification round one batched compress and decompress of the fields still
 * pending, judged exactly per field; the vertices of changed triangles fall
 * back to the analytic bound. Streams (decodable by the plain 2D decoder, with
 * the originals' critical points) are returned like the batch codec's. */
int cpsz_compress_2d_learned_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                                       const float* const* dV, const size_t* r1, const size_t* r2,
                                       double max_pwr_eb, const float* params, int tune,
                                       const unsigned char** d_base, size_t* offsets, size_t* sizes,
                                       int* iterations, size_t* boosted, size_t* fallback);
/* Learned-bound compression of a batch of 3D fields: every verification
 * round compresses the fields still pending in one batched call (one
 * persistent wavefront launch), decodes them in one batched call and judges
 * them exactly; the vertices of tetrahedra whose critical point changed fall
 * back to the analytic bound. Streams (decodable by the plain 3D decoder, with
 * the originals' critical points) are returned like the batch codec's. */
int cpsz_compress_3d_learned_batch_dev(int device, void* stream, size_t nfields, const float* const* dU,
                                       const float* const* dV, const float* const* dW, const size_t* r1,
                                       const size_t* r2, const size_t* r3, double max_pwr_eb,
                                       const float* params, int tune, const unsigned char** d_base,
                                       size_t* offsets, size_t* sizes, int* iterations, size_t* boosted,
                                       size_t* fallback);
/* Device check of the premises behind the 3D codec's approximate bound pass
 * (no reference counterpart; for tests): *max_ulp = largest ulp distance
 * between the fast and the exact cube root over every float in [0, 2^126)
 * (the pass assumes <= 2); *max_ratio = largest |fast - exact| sign bound
 * divided by its error envelope over hashed argument pairs (<= 1 when sound). */
int cpsz_check_fast_bounds(int device, uint32_t* max_ulp, double* max_ratio);
/* Process-wide choice of the 2D bounds kernel for float fields (no reference
 * counterpart; for tests and diagnostics): 0 (default) = FP32 approximations of
 * the last FP64 steps with exact fix-ups, 1 = every step in FP64. Both give
 * identical streams. */
int cpsz_set_exact_2d_bounds(int on);
/* Host evaluations of the approximate 2D bound's pieces (no reference
 * counterpart; for tests, no GPU needed): the cell-table entry of each
 * triangle uv6[6i..6i+5] = (u0,u1,u2,v0,v1,v2) with square roots and
 * reciprocals moved by up to `ulps` units in the last place, and the
 * division-free test of (det / d >= 1). */
int cpsz_tri_bound_2d_fast_host(const float* uv6, size_t n, int ulps, float* entry);
int cpsz_ratio_ge1_host(const double* det, const double* d, size_t n, unsigned char* out);
/* Host evaluation of the exact tetrahedron predicate the device judge uses:
 * f = 4 vertex vectors (u, v, w) of tetrahedron t (0..5, the judge's order);
 * *type = -1 (no critical point) or the type code. */
int cpsz_tet_cp_emulate(const float* f, int t, int* type);

/* ---- instrumentation ------------------------------------------------------ */
/* Enables per-stage HIP-event timing for this thread's context on `device`. */
int cpsz_set_timing(int device, int on);
/* Stage timings of the last call: names joined by ';' into `names`, times in
 * ms into `ms`; returns the number of stages (or a negative error). */
int cpsz_last_stages(int device, char* names, size_t names_cap, float* ms, int cap);

#ifdef __cplusplus
}
#endif

#endif /* CPSZ_H_ */
