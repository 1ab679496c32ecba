"""CPU oracle for the MCTomo likelihood forward model -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker.  The product library never calls it.
"""
