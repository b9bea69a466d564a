"""ORACLE - TEST INFRASTRUCTURE ONLY: the host restatement of the reference's
SPH pipeline, used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. The product never imports it."""
