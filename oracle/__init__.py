"""CPU oracle of the state-vector hot path -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, as the checker or as the timed CPU baseline; never by qcs_amd/.
"""
