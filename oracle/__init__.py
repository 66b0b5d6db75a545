"""CPU oracle for the Pauli-vector hot path -- TEST INFRASTRUCTURE ONLY.

A clean-room numpy float64 restatement of quantumsim's reference backend
(quantumsim/pauli_vectors/numpy.py, pauli_vector.py, bases/, algebra/).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
it, and only to check results or as the CPU baseline; the product
(quantumsim_amd) never does.

Pinning (what checks each function, see tests/test_oracle.py):
* bases.general / bases.gell_mann: the reference's own known-answer index
  tests (test_pauli_vectors.py:122-125, test_operations_qutrits.py:76-93,
  test_compiler.py:301) restated as fixtures, orthonormality, closed forms.
* algebra.kraus_to_ptm: the amplitude-damping closed form
  (test_operations.py:34-37), CPTP trace preservation, Hilbert brute force.
* algebra.dm_to_pv / pv_to_dm: from_dm coefficient KATs
  (test_pauli_vectors.py:122-125), round trip.
* PauliVectorNumpy.apply_ptm: Hilbert-space brute force rho' = K rho K^dag
  for 1-3 targets in every order, sequential == fused.
* diagonal / trace / meas_prob / partial_trace: np.diag of pv_to_dm and a
  Hilbert-space partial trace on small inputs; the reference's KATs
  (test_pauli_vectors.py:239-307, test_operations_qubits.py:13-141).
* multitake: numpy's x[np.ix_(...)].
"""
