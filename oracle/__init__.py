"""The oracle: psyref, a C++ restatement of the psychopath CPU path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by
bench.py's CPU-baseline leg, always as the checker or the baseline, never as
the product.
"""
