"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the refinement path.

Only tests/, the smoke check of ``__graft_entry__`` and the CPU-baseline leg
of ``bench.py`` may import this package, and only as the checker.
"""
