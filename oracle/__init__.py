"""ORACLE — test infrastructure only: the CPU restatement of the reference hot
path used to check the HIP implementation.  Only tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py may import it."""
