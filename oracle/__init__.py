"""CPU restatement of jump's explicit ULMPM step: the parity oracle.

Test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  The product library never links or calls it."""
