"""CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.c).  Used by tests/, by the smoke check
and by bench.py's cpu_baseline leg as the checker / baseline, never by the product."""
