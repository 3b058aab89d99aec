"""Test oracle for the thick-restart Lanczos path (test infrastructure only).

Only tests/, the smoke check and the bench's CPU-baseline leg may import this
package, and only as the checker; the product never does.  It holds
  * trlan_oracle: numpy restatement of the reference trlan() and, with
    complex_=True, ztrlan() (cited file:line per function),
  * restart_oracle / sorts: the reference's restart policy and paired sorts,
  * lapack: ctypes bindings to the LAPACK routines the reference calls,
  * ref: a ctypes wrapper of the compiled reference (oracle/_ref, built by
    oracle/Makefile from the reference sources), used to pin the restatement.
"""
