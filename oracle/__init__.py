"""CPU oracle for the TileStream .tsc codec -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package.  The product library (tilestream_amd) never imports it.
"""
