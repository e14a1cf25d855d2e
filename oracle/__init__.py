"""ORACLE -- test infrastructure only.

CPU restatement of the reference's operator path (oracle.c + glue.py) and bindings to the
reference's own compiled CPU libraries (ref.py, built into oracle/_ref by oracle/Makefile).
Only tests/, the golden-fixture script, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it, and only as the checker; the product never imports it.
"""
