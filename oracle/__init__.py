"""CPU oracle of the SBP-SAT hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, always as the checker (or the timed CPU baseline), never as a path
that produces the library's results.
"""
