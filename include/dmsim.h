/*
 * dmsim.h -- C ABI of the MI355X Pauli-basis density-matrix engine.
 *
 * The state of an n-qubit register is the vector of its 4^n real Pauli
 * coefficients a_{i_0 ... i_{n-1}} (i in {0,1,2,3} = {I,X,Y,Z}), resident in
 * HBM.  Qubit q is the digit of weight 4^(n-1-q) of the flat index (qubit 0 is
 * the slowest digit), the layout of the reference backend
 * (qiskit/providers/basicaer/dm_simulator.py:348-349).  Storage is FP64 or
 * FP32; every reduction accumulates in FP64.
 *
 * Every entry point returns 0 on success or a negative DM_ERR_* code;
 * dm_last_error() gives the message of the calling thread's last failure.
 * A handle owns one HIP stream and its device buffers; calls on one handle
 * must not overlap (separate handles may run on separate threads).  Calls
 * that return data to the host synchronise the handle's stream; the others
 * only enqueue work.
 *
 * Each declaration names the reference function it replaces.
 */
#ifndef DMSIM_H
#define DMSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dm_state dm_state;

enum { DM_FP64 = 0, DM_FP32 = 1 };

enum {
  DM_OK = 0,
  DM_ERR_ARG = -1,     /* bad argument (qubit index, size, null pointer) */
  DM_ERR_HIP = -2,     /* HIP runtime failure (incl. out of device memory) */
  DM_ERR_STATE = -3    /* handle in the wrong state for the call */
};

/* Message of the calling thread's last failure (always NUL-terminated). */
int dm_last_error(char *buf, size_t len);

/* Library / device description: number of visible devices, compute units and
 * HBM bytes of device `device`.  No counterpart in the reference. */
int dm_device_info(int device, int *n_devices, int *cu_count, uint64_t *hbm_bytes);

/* Allocate an n-qubit state (n in [1, 20]) on the current HIP device.
 * Replaces the numpy state owned by DmSimulatorPy (dm_simulator.py:143-175). */
int dm_create(int n_qubits, int dtype, dm_state **out);
int dm_destroy(dm_state *h);
int dm_num_qubits(const dm_state *h, int *n);
int dm_dtype(const dm_state *h, int *dtype);
/* Device address of the coefficient vector (for timing harnesses). */
int dm_device_ptr(const dm_state *h, void **ptr);

/* Product state: a = (x)_q v_q with v_q = per_qubit4[4q .. 4q+3] (qubit 0
 * slowest), times `scale`.  Replaces the np.kron construction of
 * _initialize_densitymatrix (dm_simulator.py:284-349). */
int dm_init_product(dm_state *h, const double *per_qubit4, double scale);

/* Host <-> device copies of all 4^n coefficients as FP64 (converted in the
 * FP32 form).  dm_upload replaces the stored-state load (dm_simulator.py:
 * 337-343); dm_download the coefficient export of _get_densitymatrix
 * (:1257-1269).  dm_download_chop zeroes |a| < thr on the device first (the
 * reference chops the live state, :1262-1263). */
int dm_upload(dm_state *h, const double *host);
int dm_download(dm_state *h, double *host);
int dm_download_chop(dm_state *h, double thr, double *host);
/* Chunked forms for states too large for one host array (checkpoints of
 * 16-17 qubit states): coefficients [offset, offset + count) to or from a
 * host buffer of host_dtype (DM_FP64 or DM_FP32; converted on the device in
 * chunks of at most 256 MB when it differs from the storage type), and the
 * FP64 dot product sum_i a_i b_i of that range with host values b.
 * Replace the whole-array np.save / np.load of _store_density_matrix
 * (dm_simulator.py:1271-1275), the stored-state load (:337-343) and the
 * overlap of the compare option (:262-271, 1277-1282). */
int dm_download_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, void *host);
int dm_upload_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, const void *host);
int dm_dot_range(dm_state *h, uint64_t offset, uint64_t count, int host_dtype, const void *host, double *sum);
/* Read k single coefficients by flat index. */
int dm_read_coeffs(dm_state *h, const uint64_t *index, int k, double *out);

/* General single-qubit Pauli transfer matrix on qubit q: rows X,Y,Z of the
 * 4x4 PTM, row-major (R[4*r + c] is the coefficient of input component c in
 * output component r+1); the I row is the identity (trace preserving).
 * Replaces rot_gate_dm_matrix via _add_unitary_single (basicaertools.py:
 * 93-126, dm_simulator.py:367-384), _add_qasm_reset (:810-823),
 * _add_qasm_measure_{X,Y,Z,N} (:574-704). */
int dm_apply_1q(dm_state *h, int q, const double R[12]);

/* CX(control, target) as a signed permutation of the 16 two-qubit Pauli
 * pairs, each output pair (c,t) (index 4c+t, control component first)
 * multiplied by scale[4c+t] (pass NULL for a noiseless CX; a per-gate
 * depolarising channel is a scale).  Replaces cx_gate_dm_matrix with
 * tsp_model_error = [1, 0] (basicaertools.py:310-392). */
int dm_apply_cx(dm_state *h, int control, int target, const double *scale16);

/* Dense 16x16 two-qubit PTM on (qa, qb): R[16*r + c], pair index 4*i_qa + i_qb.
 * Used for the transition-selective-pulse CX error (basicaertools.py:
 * 328-388) and fused two-qubit blocks. */
int dm_apply_2q(dm_state *h, int qa, int qb, const double R[256]);

/* Memory noise on every qubit: x,y <- sqrt(g) f (x,y); z <- g z + (1-g)(2p-1) a_I.
 * Replaces _add_decoherence_and_amp_decay (dm_simulator.py:397-425). */
int dm_memory_noise(dm_state *h, double f, double p, double g);

/* Per-qubit diagonal scaling: a_i *= prod_q tab[4q + i_q] (collapse onto a
 * Pauli string, multi-qubit depolarising).  Replaces the slice scaling of
 * _pauli_string_expectation (dm_simulator.py:540-572). */
int dm_apply_diag(dm_state *h, const double *tab4n);

/* Fused block: a k-qubit Pauli transfer matrix S (4^k x 4^k, FP64 host,
 * row-major, component index sum_j i_{qubits[j]} 4^j) applied to the state
 * (MFMA for k = 3, 4).  No reference counterpart (the reference applies gates
 * one by one). */
int dm_apply_block(dm_state *h, const int *qubits, int k, const double *S);

/* Outcome probabilities of measuring qubits[0..m-1] (ascending) along basis
 * 'X','Y','Z' (or 'N' with unit vector nvec) with polarisation retention err,
 * marginalised over the other qubits: out[b], b's most significant bit is
 * qubits[0].  FP64 gather + Walsh-Hadamard contraction on the device.
 * Replaces _add_ensemble_measure / _add_partial_measure's probabilities
 * (dm_simulator.py:427-538). */
int dm_probabilities(dm_state *h, const int *qubits, int m, char basis,
                     const double *nvec, double err, double *out);

/* tr = 2^n a_0; purity = 2^n sum a^2 (FP64). */
int dm_trace_purity(dm_state *h, double *tr, double *purity);

/* 2^n <a, b> with b given on the host (FP64).  Replaces _state_overlap
 * (dm_simulator.py:1277-1282). */
int dm_overlap(dm_state *h, const double *other_host, double *out);

/* Two states of the same register (any storage types) compared on the device
 * in FP64: max |a-b|, sum a*b, sum a^2, sum b^2 (state fidelity of the
 * reference's compare option is 2^n sum a*b, dm_simulator.py:1277-1282). */
int dm_compare(dm_state *a, dm_state *b, double *max_abs_diff, double *dot, double *norm2_a, double *norm2_b);

/* 2^n x 2^n complex128 density matrix (row-major, interleaved re/im), qubit 0
 * the most significant bit of row and column.  Replaces _compute_densitymatrix
 * (dm_simulator.py:1198-1255). */
int dm_to_dense(dm_state *h, double *out_c128);

/* A program: a list of operations executed back to back on the handle's
 * stream with one upload of all block matrices and no host synchronisation.
 * Replaces the per-level dispatch loop of run_experiment (dm_simulator.py:
 * 1005-1177) for the unitary / channel part of a circuit.
 * Parameters of op i start at params[ops[i].param]:
 *   DM_OP_1Q            qubits[0], 12 values (rows X,Y,Z of the 4x4 PTM)
 *   DM_OP_CX            qubits[0]=control, qubits[1]=target, no values
 *   DM_OP_CX_SCALED     as DM_OP_CX, 16 per-pair scales
 *   DM_OP_2Q            qubits[0..1], 256 values (dense 16x16 PTM)
 *   DM_OP_BLOCK         k = 3 or 4 qubits, 4^(2k) values (block PTM)
 *   DM_OP_MEMORY_NOISE  3 values f, p, g
 *   DM_OP_DIAG          4n values (per-qubit diagonal tables)
 *   DM_OP_FUSED         one HBM pass applying several operations to tiles
 *                       of 4^m coefficients staged in LDS; a variable-length
 *                       record of values (integers stored exactly):
 *                         m (5, 6 or 7), nq, q_1 .. q_nq  -- qubits the tile
 *                           holds; it is completed with the qubits of lowest
 *                           stride, and must hold qubit n-1
 *                         nsub, then nsub sub-operations, each
 *                           kind, kq, q0, q1, q2, q3, nv, v_1 .. v_nv
 *                         kind 1: 1-qubit PTM rows (kq=1, nv=12)
 *                         kind 2: CX control q0, target q1 (kq=2, nv=0 or
 *                                 16 per-pair scales)
 *                         kind 3: dense 16x16 PTM on (q0,q1) (kq=2, nv=256)
 *                         kind 4: dense 4^kq x 4^kq PTM on q0..q{kq-1}
 *                                 (kq=3,4; m=6; nv=4^(2kq); matrix cores)  */
enum {
  DM_OP_1Q = 1,
  DM_OP_CX = 2,
  DM_OP_CX_SCALED = 3,
  DM_OP_2Q = 4,
  DM_OP_BLOCK = 5,
  DM_OP_MEMORY_NOISE = 6,
  DM_OP_DIAG = 7,
  DM_OP_FUSED = 8
};

typedef struct dm_op {
  int32_t kind;
  int32_t k;          /* number of entries of qubits[] in use */
  int32_t qubits[4];
  int64_t param;      /* offset of the op's values in params[] */
} dm_op;

int dm_run_program(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams);

/* dm_run_program without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The program image (tile
 * operation records and block / fused-pass values, in the storage type) goes
 * to a fixed 8 MiB device array that every call on the same GPU shares and
 * that keeps the last call's contents; larger images are refused with
 * DM_ERR_ARG before any work.  Calls that overlap on one GPU (other handles,
 * other threads) race on that array.  Same results as dm_run_program (the
 * same kernels). */
int dm_run_program_noalloc(dm_state *h, const dm_op *ops, int nops, const double *params, int64_t nparams);

/* Fusion cost model (no reference counterpart): a 3-layer perceptron
 * x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3, FP32, whose forward,
 * backward (mean squared error) and Adam kernels run on the GPU.  Parameter
 * vector: W1 [F x H] row-major, b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1].
 * dm_mlp_param_count returns the length (or a negative error code). */
int dm_mlp_param_count(int n_features, int hidden);
int dm_mlp_init(int n_features, int hidden, uint64_t seed, float *params);
int dm_mlp_predict(const float *params, int n_features, int hidden, const float *X, int n, float *out);
int dm_mlp_loss_grad(const float *params, int n_features, int hidden, const float *X, const float *y, int n,
                     float *loss, float *grad);
int dm_mlp_train(float *params, int n_features, int hidden, const float *X, const float *y, int n, int epochs,
                 int batch, float lr, uint64_t seed, float *loss_per_epoch);

/* Batches of independent circuits (no reference counterpart; the reference
 * runs experiments one after another, dm_simulator.py:935-936).
 * Circuit c has n_qubits[c] (1..12) qubits and the program ops[op_start[c] ..
 * op_start[c+1]) (dm_op records over the shared params array).  Each circuit
 * starts in |0...0>, runs its program and ends with the Z-basis ensemble
 * probabilities of all its qubits with polarisation retention meas_err[c]
 * (2^n values, most significant bit = qubit 0, circuits concatenated).
 * Circuits of at most lds_max_qubits (5..7) qubits run whole inside one
 * workgroup's LDS (one launch per width, lds_threads threads per workgroup).
 * The others keep their state in one HBM arena: each operation becomes a
 * 6-digit tile pass (DM_OP_FUSED with tile size 6, DM_OP_1Q, DM_OP_CX[_SCALED],
 * DM_OP_2Q, DM_OP_MEMORY_NOISE; memory noise as 1-qubit channels, 6 qubits
 * per pass; in FP32 also DM_OP_FUSED with tile size 7 holding single-qubit
 * and CX operations).  Round r holds pass r of every circuit, plus the
 * |0..0> initialisation first and the probabilities last: its
 * initialisation and probability work (split into at most items_per_pass
 * items per pass) is one grid-stride launch, and its 6-digit and 7-digit
 * tile passes one launch each with a workgroup per tile, the three side by
 * side.  A run is at most 7 LDS-tier launches plus three launches per round
 * (the longest circuit's passes + 2 rounds); dm_batch_run with use_graph != 0
 * captures them into one hipGraph on first use and replays it.  Per-circuit
 * times (circuit_us) are GPU clock stamps from a circuit's initialisation to
 * its probabilities; circuits run concurrently.  dm_batch_info reports the
 * HBM-tier circuit, pass and item counts and the launches of one run. */
typedef struct dm_batch dm_batch;
int dm_batch_create(int dtype, int ncirc, const int *n_qubits, const int64_t *op_start, const dm_op *ops,
                    const double *params, int64_t nparams, const double *meas_err, int lds_max_qubits,
                    int lds_threads, int items_per_pass, dm_batch **out);
int dm_batch_info(const dm_batch *b, int64_t *hbm_circuits, int64_t *passes, int64_t *items, int64_t *launches);
int dm_batch_run(dm_batch *b, int use_graph, double *probs_out, double *circuit_us, double *total_ms);
int dm_batch_destroy(dm_batch *b);

/* The program of one batch circuit, built on the host (no device work; no
 * reference counterpart: the reference interprets each experiment's
 * instructions level by level, dm_simulator.py:1005-1177, with a barrier
 * after every gate giving one memory-noise level per gate, :1172-1177).
 * Gate i: kind[i] = DM_INST_U1 (angles[3i] = lambda), DM_INST_U3 (angles
 * [3i..3i+2] = theta, phi, lambda; the PTM rz(phi) ry(theta) rz(lambda),
 * basicaertools.py:69-90) or DM_INST_CX (q0 control, q1 target); every gate
 * is followed by the depolarising channel gate_depolarization on its qubits
 * (none at 0) and by memory noise (decoherence, thermal, decay factors, the
 * reference's retention factors).  fused == 0: each gate as its DM_OP_1Q /
 * DM_OP_CX[_SCALED] record, then one DM_OP_MEMORY_NOISE.  fused != 0 (n >= 6):
 * memory noise as one 1-qubit channel per qubit (none when decay and
 * decoherence are 1), the greedy plan of dm_plan_fusion (blocks of up to 4
 * qubits), and each block as one DM_OP_FUSED pass of 6 digits (FP32: 7 where
 * 6 would leave contiguous runs of 4^2 coefficients that 7 lengthens, or
 * runs of 4^3 that 7 makes 4^4), runs of 1-qubit PTMs on one qubit folded;
 * single-gate blocks keep their plain records and one-qubit blocks become
 * DM_OP_1Q.  precision (DM_FP64 / DM_FP32) is the state's.  *nops and
 * *nvalues receive the sizes; with ops and values both NULL only the sizes
 * are computed, else the records go to ops[0..max_ops) and values
 * [0..max_values) (DM_ERR_ARG when they do not fit). */
int dm_batch_circuit_program(int n, int ngates, const int32_t *kind, const int32_t *q0, const int32_t *q1,
                             const double *angles, double decay_factor, double decoherence_factor,
                             double thermal_factor, double gate_depolarization, int fused, int precision, dm_op *ops,
                             int max_ops, double *values, int64_t max_values, int *nops, int64_t *nvalues);

/* Host front end (no device work; plain C++ over integer arrays).
 *
 * Instruction kinds of the merge and the partition: */
enum {
  DM_INST_U1 = 1,          /* u1(lambda): params[0] */
  DM_INST_U3 = 2,          /* u3(theta, phi, lambda); u2 and U arrive as u3 */
  DM_INST_ID = 3,          /* id / u0: dropped by the merge */
  DM_INST_SEPARATOR = 4,   /* cx, measure, bfunc, reset, barrier: flush pending runs */
  DM_INST_ROT = 5,         /* partition: u1 / u3 on q0 */
  DM_INST_CX = 6,          /* partition: cx on (q0, q1) */
  DM_INST_MEASURE = 7,     /* partition: measure of q0 */
  DM_INST_RESET = 8,       /* partition: reset of q0 */
  DM_INST_BARRIER = 9,     /* partition: segment boundary */
  DM_INST_OWN_LEVEL = 10,  /* partition: a Bell / Expect / Ensemble measure; at the head of a
                              segment the whole segment is one level, elsewhere a measure */
  DM_INST_HOLD_MEASURE = 11, DM_INST_HOLD_RESET = 12  /* internal placeholders */
};

/* Single-qubit gate merging: single_gate_merge with merge_gates / mergeU /
 * U3_merge (basicaertools.py:129-307).  With merge != 0 every run of u1/u3
 * on one qubit between separators folds into one gate that reuses the run's
 * first record, and every separator flushes all pending runs in qubit order;
 * with merge == 0 the rotations pass through.  id / u0 are dropped.  Output
 * record j: out_src[j] (the input record it reuses), out_kind[j] (U1, U3 or
 * SEPARATOR), out_params[3j..3j+2].  Angles are evaluated with the C
 * library's sin, cos, atan2, acos (the reference uses numpy's). */
int dm_merge_rotations(int ninst, const int32_t *kind, const int32_t *qubit, const double *params, int merge,
                       int32_t *out_src, int32_t *out_kind, double *out_params, int *nout);

/* Level partition: qubit_stack / partition_helper / partition
 * (basicaertools.py:536-748).  Barriers split segments; each segment's
 * per-qubit stacks are consumed greedily (a cx when it heads both stacks; all
 * measures (resets) together once every stack head is one or its
 * placeholder).  Output: the levels of all segments in order, each padded to
 * its segment's deepest stack as the reference's table is:
 * out_ops[out_level_start[l] .. out_level_start[l+1]) are the instruction
 * indices of level l (nlevels levels); *count is the number of levels the
 * experiment loop runs.  out_ops holds ninst entries, out_level_start
 * 4 * ninst + 4 (each pass of the level loop places an instruction or
 * stalls, and three stalls in a row are an error). */
int dm_partition(int n, int ninst, const int32_t *kind, const int32_t *q0, const int32_t *q1, int32_t *out_ops,
                 int32_t *out_level_start, int *nlevels, int *count);

/* Fusion planner (no reference counterpart): gates are noisy 1- or 2-qubit
 * PTMs, gate_qubits[2i..2i+1] (second -1 for one qubit).  A block is a set
 * of gates that can run together: its gates keep their order on every qubit
 * and every gate left out of it that comes earlier acts on other qubits.
 * method: DM_PLAN_GREEDY (grow each block to kmax qubits), DM_PLAN_SEARCH
 * (the candidate with the most gates), DM_PLAN_LEARNED (candidates scored by
 * their predicted net saving under the cost model mlp_params, with the
 * lookahead widths lookahead[0..nlookahead)); tiles != 0 adds the tile-wide
 * candidates of the learned planner.  precision (DM_FP64 / DM_FP32) is the
 * state's: it selects the cost model's precision and the tile sizes (FP32
 * runs 3 qubits above the three lowest positions in 7-digit tiles, FP64 in
 * 6-digit ones).  Output: blocks in execution order,
 * out_members[out_block_start[b] ..] (gate indices) and
 * out_qubits[out_qubit_start[b] ..] (qubits in order of first use); out
 * arrays hold ngates, ngates + 1, 2 * ngates and ngates + 1 entries. */
enum { DM_PLAN_GREEDY = 0, DM_PLAN_SEARCH = 1, DM_PLAN_LEARNED = 2 };
int dm_plan_fusion(int n, int ngates, const int32_t *gate_qubits, int method, int kmax, int tiles,
                   const int32_t *lookahead, int nlookahead, const float *mlp_params, int precision,
                   int32_t *out_members, int32_t *out_block_start, int32_t *out_qubits, int32_t *out_qubit_start,
                   int *nblocks);
/* The cost model's forward pass on the host (FP32, fixed summation order;
 * the planner's predictions) and its predicted seconds for given blocks
 * (the same arrays as dm_plan_fusion's output, precision DM_FP64/DM_FP32). */
int dm_mlp_predict_host(const float *params, int n_features, int hidden, const float *X, int n, float *out);
int dm_cost_block_seconds(int n, int ngates, const int32_t *gate_qubits, int nblocks, const int32_t *members,
                          const int32_t *block_start, const int32_t *qubits, const int32_t *qubit_start,
                          const float *mlp_params, int precision, double *out);

/* Stream control and timing (no reference counterpart). */
int dm_synchronize(dm_state *h);
/* Accumulated GPU time per kernel family since the last reset, measured with
 * HIP events around each launch when profiling is on. family: 0 = 1q, 1 = cx,
 * 2 = 2q dense, 3 = memory noise, 4 = diag, 5 = block k=3, 6 = reductions,
 * 7 = init/copies, 8 = block k=4, 9 = fused tile passes. */
int dm_profile(dm_state *h, int enable);
int dm_profile_read(dm_state *h, int family, uint64_t *launches, double *ms);
int dm_profile_reset(dm_state *h);

#ifdef __cplusplus
}
#endif

#endif /* DMSIM_H */
