"""CPU transcription of SenseMap's PatchMatch stage -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it, as the
checker and as the CPU baseline; the product never imports it.
"""
