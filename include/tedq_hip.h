/*
 * tedq_hip.h — C ABI of the MI355X (gfx950) execution engine for TeD-Q compiled circuits.
 *
 * The library (ted-q_amd/tedq_amd/_lib/libtedq_hip.so) is bound from Python
 * with ctypes (ted-q_amd/tedq_amd/engine/hip.py).  Conventions for every entry point:
 *   - pointers named dev_* / state / psi / lam / mats / partials / out are DEVICE
 *     pointers allocated by the caller (torch's caching allocator); the library
 *     never allocates or frees device memory;
 *   - pointers documented as "host" are read on the host during the call;
 *   - `stream` is a hipStream_t (the caller's current stream); all work is
 *     enqueued asynchronously and no entry point synchronises;
 *   - the return value is 0 on success, negative on error; tedq_last_error()
 *     returns the message of the last failure on the calling thread.
 * State layout: a batch of circuits is [batch][2^n] complex64 (interleaved
 * float re/im), qubit q of a circuit is flat bit n-1-q (the reference's
 * C-order [2]*n tensor, tedq/backends/pytorch_backend.py:513-522).
 */
#ifndef TEDQ_HIP_H
#define TEDQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library plumbing ---------------------------------------------------- */
const char* tedq_last_error(void);
int tedq_abi_version(void);
/* props[0..3] = {CU count, max LDS bytes per workgroup, wavefront size, arch major*100+minor} */
int tedq_device_info(int* props /* host, 4 ints */);

/* ---- state preparation --------------------------------------------------- */
/* |0...0> for every circuit.  Replaces default_initstate
 * (tedq/backends/pytorch_backend.py:513-522). */
int tedq_sv_init_zero(void* state, int64_t batch, int n_qubits, void* stream);
/* Broadcast one user state (dev, 2^n complex64) to every circuit.  Replaces
 * get_initstate with an IintStateVector (pytorch_backend.py:500-510). */
int tedq_sv_init_copy(void* state, const void* init, int64_t batch, int n_qubits, void* stream);

/* ---- gate tables ----------------------------------------------------------- */
/* Per (circuit, slot): fused gate matrix (complex64) and, when amats != NULL,
 * per trainable parameter the fp64 matrix A with dL/dtheta = 2 Re sum A_ij C_ij.
 * params: dev [batch][n_params] float32 (params_are_f64 = 0) or float64.
 * slot_recs / comp_recs: dev int32 records (8 words each, csrc/sv_tables.hip);
 * const_pool: dev complex128 constants; err_flag: dev int, OR-ed with 1 on an
 * unknown gate code.  Replaces the per-gate tensor builders
 * (pytorch_backend.py:579-1188) and _parser_circuit (compiled_circuit.py:418-440). */
int tedq_sv_build_tables(const void* params, int params_are_f64, int64_t batch, int n_params,
                         const void* slot_recs, int n_slots, const void* comp_recs, const void* const_pool,
                         void* mats, int64_t mat_stride, void* amats, int* err_flag, void* stream);

/* ---- fused passes ---------------------------------------------------------- */
/* One forward pass (a tile of 2^tile_bits amplitudes per workgroup, register
 * stages of 2^reg_bits amplitudes per thread) over every circuit; dev_prog is
 * the pass program (int32 words, csrc/sv_passes.hip), tile_mask its physical
 * tile bits.  zero_mask: non-tile bits on which the state is known to vanish
 * (a circuit started in |0..0> whose earlier passes never touched them);
 * tiles with any of these bits set are skipped and keep their (zero)
 * contents; 0 processes every tile.  A sequence of passes replaces the SV
 * propagation loop tensordot + permute per gate (pytorch_backend.py:358-380). */
int tedq_sv_apply_pass(void* state, int64_t batch, int n_qubits, int tile_bits, int reg_bits, int threads_log2,
                       const int* dev_prog, uint64_t tile_mask, uint64_t zero_mask, const void* mats,
                       int64_t mat_stride, void* stream);
/* One adjoint pass: applies the daggered gates of a pass in reverse order to
 * psi and lam and writes the gradient bilinears of each parametrised slot to
 * partials [batch][n_gslots][partial_rows][8] (float).  zero_mask as for
 * tedq_sv_apply_pass (psi vanishes there during this pass, and no later pass
 * reads lam there): skipped tiles write no partial rows (rows from
 * 2^popcount(rest bits outside zero_mask) x rows-per-tile on stay as the
 * caller set them).  Replaces torch autograd through the SV loop (the
 * default back_prop method, pytorch_backend.py:142-144). */
int tedq_sv_adjoint_pass(void* psi, void* lam, int64_t batch, int n_qubits, int tile_bits, int reg_bits,
                         int threads_log2, const int* dev_prog, uint64_t tile_mask, uint64_t zero_mask,
                         const void* mats, int64_t mat_stride, float* partials, int64_t partial_rows, int n_gslots,
                         void* stream);
/* One pass for circuits of DIFFERENT structures and sizes in one launch (a
 * serving batch), forward (adjoint = 0) or adjoint.  dev_recs: device array of
 * n_recs 64-byte records {psi, lam, prog, mats, partials (device pointers of
 * the circuit: state, adjoint state, its program of this pass, its slot-matrix
 * row, its partial rows [n_gslots][rows][8]), uint64 rest_mask (the circuit's
 * non-tile bits in this pass), int64 tile_end (tiles of this and all earlier
 * records, 2^(n-k) per circuit), int32 n, int32 n_gslots}; all circuits share
 * tile bits k and register bits r; total_tiles = the last tile_end.  Each
 * circuit is processed exactly as by tedq_sv_apply_pass / tedq_sv_adjoint_pass
 * with batch 1.  Replaces, for a batch of different circuits, the same
 * reference SV loop (pytorch_backend.py:358-380) and its autograd backward. */
int tedq_sv_pass_ragged(int adjoint, const void* dev_recs, int n_recs, int64_t total_tiles, int k, int r,
                        void* stream);
/* grad[b][p] = 2 Re sum_ij A[b][p][ij] C[b][g(p)][ij] with C reduced in fp64
 * (compensated) from the partial rows; grad is float32 or float64
 * (grad_is_f64).  pgslot: dev int32 [n_params]; C: dev double scratch
 * [batch][n_gslots][8]. */
int tedq_sv_grad_reduce(const float* partials, int64_t batch, int n_gslots, int64_t partial_rows,
                        const void* amats, const int* pgslot, int n_params, double* C, void* grad,
                        int grad_is_f64, void* stream);

/* ---- measurements ------------------------------------------------------------- */
/* Number of reduction chunks (rows of every `partial` buffer below) for n qubits. */
int64_t tedq_sv_reduce_chunks(int n_qubits);
/* out[b][m] = <psi_b| Z-string(mask_m) |psi_b> in fp64, all masks in one pass,
 * chunks combined with Neumaier summation; mask 0 is the norm.  n_masks <= 32.
 * partial: dev double [batch][chunks][n_masks].  Replaces the expectation
 * branch of get_measurement_results for PauliZ/I observables
 * (pytorch_backend.py:399-459). */
int tedq_sv_expval_z(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks, int n_masks,
                     double* partial, double* out, void* stream);
/* tedq_sv_expval_z without the partial buffer: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums go to a fixed 8 MiB device array inside the library (one per
 * GPU, held until the process ends), so batch * tedq_sv_reduce_chunks(n) *
 * n_masks * 8 must be <= 8 MiB (else -1 with a message, nothing launched).
 * Allocates nothing.  Calls that overlap on one GPU share the array and race. */
int tedq_sv_expval_z_noalloc(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks,
                             int n_masks, double* out, void* stream);
/* lam = (sum_m dy[b][m] Z-string_m) psi; n_masks may be 0 (lam = 0).  dy: dev double [batch][n_masks]. */
int tedq_sv_lambda_z(const void* psi, void* lam, int64_t batch, int n_qubits, const uint64_t* dev_masks,
                     int n_masks, const double* dy, void* stream);
/* Re <psi| Zsign(zmask) (F_0 x ... x F_{q-1}) |psi> for q <= 4 single-qubit
 * factors; bits: host int32[q] physical positions, factors: host float
 * [q][2][2][re,im].  Replaces the non-diagonal expectation branch
 * (pytorch_backend.py:402-424, 428-459). */
int tedq_sv_expval_factored(const void* state, int64_t batch, int n_qubits, int q, const int* bits,
                            const float* factors, uint64_t zmask, double* partial, double* out, void* stream);
/* lam += coef[b] * Zsign * ((F + F^dagger)/2) psi  (coef: dev double [batch]). */
int tedq_sv_lambda_factored(const void* psi, void* lam, int64_t batch, int n_qubits, int q, const int* bits,
                            const float* factors, uint64_t zmask, const double* coef, void* stream);
/* Probabilities: bits == NULL and k == n gives |psi|^2; otherwise marginals over
 * the k <= 30 physical bits listed (host int32, distinct, most significant
 * output bit first), summed in fp64 in a fixed order (no atomics).  out: dev
 * float [batch][2^k]; partial: dev double [batch][tedq_sv_probs_chunks(n, k)][2^k].
 * Replaces the probs branch (pytorch_backend.py:461-472). */
int tedq_sv_probs(const void* state, int64_t batch, int n_qubits, int k, const int* bits, double* partial,
                  float* out, void* stream);
/* Rows of the partial buffer of a k-bit marginal over n qubits (-1 for bad sizes). */
int64_t tedq_sv_probs_chunks(int n_qubits, int k);
/* Post-measurement states of ONE circuit: for every outcome y of the k listed
 * physical bits (host int32, most significant first, as for tedq_sv_probs),
 * out[y][r] = psi[x(y, r)] / sqrt(probs[y]) with r over the remaining bits in
 * C order of the remaining qubits.  probs: dev float [2^k], the marginal of the
 * same bits; out: dev complex64 [2^n].  Replaces the after_state branch of
 * probs(qubits, after_state=True) (pytorch_backend.py:474-493), from the
 * forward's state instead of a second evaluation. */
int tedq_sv_post_measure(const void* psi, int n_qubits, int k, const int* bits, const float* probs, void* out,
                         void* stream);
/* lam += g[b][marg(x)] psi_x (gradient seed of a probability output; g: dev float). */
int tedq_sv_lambda_probs(const void* psi, void* lam, int64_t batch, int n_qubits, int k, const int* bits,
                         const float* g, void* stream);
/* lam += scale * G (gradient seed of a state output; G: dev complex64). */
int tedq_sv_axpy(void* lam, const void* G, int64_t batch, int n_qubits, float scale, void* stream);
/* out[b][m] = compensated sum over chunks of partial[b][chunk][m]. */
int tedq_sum_partials(const double* partial, int64_t batch, int64_t nchunk, int M, double* out, void* stream);

/* ---- run-time specialised pass kernels ------------------------------------- */
/* Compile generated pass kernels (one per pass, ops and offsets as constants;
 * tedq_amd/engine/jit.py) with hiprtc for gfx950 and load them as a module on
 * the current device.  header_srcs/header_names: the texts and include names
 * of tedq_common.h and sv_ops.h.  Allocates host memory and device code
 * memory (the module), no data buffers.  Same pass semantics as
 * tedq_sv_apply_pass / tedq_sv_adjoint_pass, i.e. it replaces the same
 * reference SV loop (pytorch_backend.py:358-380). */
int tedq_jit_compile(const char* src, int n_headers, const char* const* header_srcs,
                     const char* const* header_names, void** module_out);
/* Kernel handle of `name` in a compiled module. */
int tedq_jit_function(void* module, const char* name, void** fn_out);
/* Per-thread scratch (private segment) bytes of a compiled kernel: nonzero when
 * the compiler spilled registers, so the caller can compile that pass in
 * another form.  No reference counterpart (a property of generated code). */
int tedq_jit_function_scratch(void* fn, int* bytes_out);
/* Unload a compiled module. */
int tedq_jit_unload(void* module);
/* Launch one generated pass (forward when adjoint == 0) over every circuit;
 * arguments as tedq_sv_apply_pass / tedq_sv_adjoint_pass (k < n <= 28, at most
 * 160 KiB of LDS per workgroup).  seed: dev double [batch][M], read only by a
 * generated first adjoint pass that forms lambda from psi (the Z-string
 * adjoint seed, tedq_sv_lambda_z, fused into the pass); null otherwise.
 * flags: 1 (forward) reads psi and writes the result to lam, a separate state
 * buffer (the forward keeps every pass's output as a checkpoint); 2 (adjoint)
 * psi is the checkpoint of the forward state at the end of this pass and is
 * only read (the state at its start is the previous checkpoint); 4 (forward,
 * with the zero mask of a sweep from |0..0>) takes the amplitudes at tile bits
 * no earlier pass held as zero instead of reading them, and the first pass
 * builds the |0..0> tile without reading psi, so psi needs no initialisation
 * when the last pass writes every tile; 8 (adjoint, kernels generated to loop
 * over tiles) launches only as many workgroups as the device holds at once,
 * each loading its next tile while it processes the current one.  0: in place. */
int tedq_jit_pass(void* fn, int adjoint, void* psi, void* lam, int64_t batch, int n_qubits, int tile_bits,
                  int reg_bits, uint64_t tile_mask, uint64_t zero_mask, const void* mats, int64_t mat_stride,
                  void* partials, int64_t partial_rows, int n_gslots, const double* seed, int flags, void* stream);

/* ---- tensor-network executor ----------------------------------------------- */
/* desc: 4 + 9*32 int32 — bit counts of the index roles (h kept+shared, m kept
 * only in A, n kept only in B, k summed), then for each of A_h A_m A_k B_h B_k
 * B_n C_h C_m C_n (32 slots each) the address bit of every role bit.  Every
 * index has dimension 2, so this fixes both operand layouts and the result
 * layout; no permuted copy is made.  Ranks are at most 30, so the last word
 * (slot 31 of C_n) is free and selects the kernel form of a GEMM-shaped step:
 * 0 = four real MFMA products per complex product on 128 x 128 tiles where
 * both free sizes allow it, 3 = the three-product (Gauss) form, 64 = four
 * products on 64 x 64 tiles only; any other value is rejected.  Bytes of device workspace the step needs
 * for its split-k partial sums (0 if none), or -1 for an invalid descriptor. */
int64_t tedq_tn_workspace_bytes(const int32_t* desc, int64_t batch);
/* One pairwise contraction step for every batch element:
 *   C[h,m,n] = sum_k A[h,m,k] B[h,k,n]   (complex64, fp32 accumulation)
 * a/b/c: dev complex64 with per-batch strides in elements (0 = shared operand);
 * desc: host copy (validated, sizes the launch) and desc_dev: an identical
 * device copy read by the kernels.  Replaces one pairwise step of the
 * reference's contraction tree.contract(arrays, backend='torch')
 * (pytorch_backend.py:276-354, oe_wrapper.py:57). */
int tedq_tn_contract(const void* a, int64_t a_bstride, const void* b, int64_t b_bstride, void* c,
                     int64_t c_bstride, int64_t batch, const int32_t* desc, const int32_t* desc_dev,
                     void* workspace, int64_t workspace_bytes, void* stream);
/* The small steps of one level of a contraction tree in one launch (steps whose
 * operands are leaves or results of earlier small steps: the gate merges that
 * make up most of a path).  bases/strides/lens: three operand arenas as device
 * pointers, per-batch strides in elements (0 = shared) and valid elements per
 * batch row; every result goes to arena 2.  records (host) / records_dev
 * (device): n_records x 8 int64 {desc row, A arena, A offset, B arena,
 * B offset, C offset, output count, 0}; descs / descs_dev: the descriptor table
 * (n_descs rows of 4 + 9*32 int32, layout as tedq_tn_contract).  Operands and
 * results of rank <= 12, at most 2^6 summed terms, bounds checked on the host.
 * Replaces the same pairwise steps of tree.contract(arrays, backend='torch')
 * (pytorch_backend.py:276-354) as tedq_tn_contract, issued together. */
int tedq_tn_contract_small_batch(const void* const* bases, const int64_t* strides, const int64_t* lens,
                                 const int64_t* records, const int64_t* records_dev, int n_records, int64_t batch,
                                 const int32_t* descs, const int32_t* descs_dev, int n_descs, void* stream);
/* Both environments of a gate step in one pass.  desc: the forward step C = A B
 * (host + device copies as in tedq_tn_contract) with one operand X large (at
 * least 2^10 rows) and the other G gate-sized (at most two summed and two new
 * indices, no kept-shared ones).  From ec = E_C (C's layout) it writes
 * ex = E_X[p,k] = sum_q E_C[p,q] G[q,k] in X's layout and eg = E_G[q,k] =
 * sum_p E_C[p,q] X[p,k] in G's layout, reading E_C and X once.  The workspace
 * query returns the bytes of partial sums needed, or -1 when the step does not
 * qualify.  Replaces the two environment contractions torch autograd performs
 * for such a step of tree.contract (pytorch_backend.py:276-354). */
int64_t tedq_tn_gate_backward_workspace(const int32_t* desc, int64_t batch);
int tedq_tn_gate_backward(const void* ec, int64_t ec_bstride, const void* x, int64_t x_bstride, const void* g,
                          int64_t g_bstride, void* ex, int64_t ex_bstride, void* eg, int64_t eg_bstride,
                          int64_t batch, const int32_t* desc, const int32_t* desc_dev, void* workspace,
                          int64_t workspace_bytes, void* stream);
/* Gate tensors of the network's leaves for every parameter row: ket U with
 * axes (out..., in...) or bra U^dagger, and dV/dtheta 2x2 blocks of trainable
 * parameters.  recs: dev int32[16] records {code, src0..2, const_off, nq,
 * kind (0 controlled 2x2 block, 1 constant matrix), bra, off, d0..2, dagger, pad};
 * leaves: dev complex64 [batch][leaf_stride]; dmats: dev complex64
 * [batch][dmat_stride].  Replaces the gate operands and their conjugates
 * (pytorch_backend.py:524-546, 567-1188) for the contraction path. */
int tedq_tn_gate_tensors(const void* params, int params_are_f64, int64_t batch, int n_params, const void* recs,
                         int nrec, const void* const_pool, void* leaves, int64_t leaf_stride, void* dmats,
                         int64_t dmat_stride, int* err_flag, void* stream);
/* grad[b][p] += upstream[b] * Re(sum E_ket . dU + sum E_bra . dU^dagger) over
 * the gate of trainable parameter p, from the leaf environments env (same
 * layout as the leaves).  recs: dev int32[4] {ket_off, bra_off, nq, dslot}.
 * With the environments from the reverse contraction tree this is the
 * gradient the reference gets from torch autograd through the contraction
 * (pytorch_backend.py:276-354). */
int tedq_tn_param_grad(const void* env, int64_t env_stride, const void* dmats, int64_t dmat_stride,
                       const void* recs, int n_params, const float* upstream, double* grad, int64_t batch,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TEDQ_HIP_H */
