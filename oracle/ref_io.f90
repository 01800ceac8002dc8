! Test infrastructure: closes a Fortran unit the reference wrote to, so that its
! buffered records reach the file before the oracle reads it back.
SUBROUTINE xt_ref_close(unit)
IMPLICIT NONE
INTEGER, INTENT(IN) :: unit
CLOSE(unit)
END SUBROUTINE xt_ref_close
