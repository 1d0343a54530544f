C=======================================================================
C     wd2_shim.f -- Fortran side of the wolfd2_amd drop-in boundary.
C
C     Exports the reference's hot-path entry points under their own names
C     and argument lists (F77 implicit interfaces, every argument by
C     reference), so a wolfd2 driver links against them unchanged:
C
C       nAuxMomentum   momentum.f:33-48
C       Ppe            pressure.f:30-38
C       Project        utility.f:253-259
C       VelBoundCond   bound_cond.f:511-513
C       PresBoundCond  bound_cond.f:853-856
C       VelOutflowBCs  bound_cond.f:1656-1659
C       DiffMaxNorm    utility.f:446
C       DMaxNorm       utility.f:479
C       TempBoundCond  bound_cond.f:1030-1033
C       ThermEnergy    thermal.f:24-32
C       EqState        thermal.f:280-282
C       Filter         utility.f:31-34
C
C     Each one calls the C ABI of libwolfd2_amd (include/wolfd2_amd.h)
C     through ISO_C_BINDING interfaces.  The caller's compile-time extents
C     come from its own config.f (mnx, mgri, mgrj): every field is
C     (0:mnx,0:mny), so its leading dimension is mnx+1.  A fatal status
C     (bad boundary or solver flag, device failure) prints the library's
C     message and STOPs, as the reference does for a bad flag.
C
C     Fixed form so that the caller's config.f can be INCLUDEd as is.
C     Build: amdflang -ffixed-form -I<dir of config.f> -c wd2_shim.f
C=======================================================================
      MODULE wd2_iface
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTERFACE
        FUNCTION wd2_last_error() BIND(C, NAME='wd2_last_error')
          IMPORT :: C_PTR
          TYPE(C_PTR) :: wd2_last_error
        END FUNCTION

        FUNCTION wd2_naux_momentum(nx, ny, ld, mgri, mgrj, mqiter,
     &      nReg, nRegBrd, nRegType, nMomBdTp, dk, re, fr, qtol,
     &      dPRporos, dPRporc1, dPRporc2, dBCVal,
     &      ran, rbn, rgn, rac, rbc, rgc, dju, djv,
     &      xec, yec, xzn, yzn, xen, yen, xzc, yzc,
     &      xeu, yeu, xzu, yzu, xev, yev, xzv, yzv,
     &      d, dn, un, vn, us, vs, nQLiter)
     &      BIND(C, NAME='wd2_naux_momentum')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_naux_momentum
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj, mqiter
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nRegType(*)
          INTEGER(C_INT) :: nMomBdTp(*), nQLiter
          REAL(C_DOUBLE), VALUE :: dk, re, fr, qtol
          REAL(C_DOUBLE) :: dPRporos(*), dPRporc1(*), dPRporc2(*)
          REAL(C_DOUBLE) :: dBCVal(*)
          REAL(C_DOUBLE) :: ran(*), rbn(*), rgn(*), rac(*), rbc(*)
          REAL(C_DOUBLE) :: rgc(*), dju(*), djv(*)
          REAL(C_DOUBLE) :: xec(*), yec(*), xzn(*), yzn(*)
          REAL(C_DOUBLE) :: xen(*), yen(*), xzc(*), yzc(*)
          REAL(C_DOUBLE) :: xeu(*), yeu(*), xzu(*), yzu(*)
          REAL(C_DOUBLE) :: xev(*), yev(*), xzv(*), yzv(*)
          REAL(C_DOUBLE) :: d(*), dn(*), un(*), vn(*), us(*), vs(*)
        END FUNCTION

        FUNCTION wd2_ppe(nx, ny, ld, mgri, mgrj, nReg, nRegBrd,
     &      nRegType, lCartesGrid, nPpeSolver, msorit, nSorConv,
     &      dk, sortol, sorrel, rau, rbu, rbv, rgv,
     &      xeu, yeu, xzv, yzv, u, v, p)
     &      BIND(C, NAME='wd2_ppe')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_ppe
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT), VALUE :: lCartesGrid, nPpeSolver, msorit
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nRegType(*)
          INTEGER(C_INT) :: nSorConv
          REAL(C_DOUBLE), VALUE :: dk, sortol, sorrel
          REAL(C_DOUBLE) :: rau(*), rbu(*), rbv(*), rgv(*)
          REAL(C_DOUBLE) :: xeu(*), yeu(*), xzv(*), yzv(*)
          REAL(C_DOUBLE) :: u(*), v(*), p(*)
        END FUNCTION

        FUNCTION wd2_project(nx, ny, ld, mgri, mgrj, nReg, nRegBrd,
     &      nRegType, nMomBdTp, dk, dju, djv, yeu, xzv, yzu, xev,
     &      p, u, v) BIND(C, NAME='wd2_project')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_project
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nRegType(*)
          INTEGER(C_INT) :: nMomBdTp(*)
          REAL(C_DOUBLE), VALUE :: dk
          REAL(C_DOUBLE) :: dju(*), djv(*), yeu(*), xzv(*)
          REAL(C_DOUBLE) :: yzu(*), xev(*), p(*), u(*), v(*)
        END FUNCTION

        FUNCTION wd2_vel_bound_cond(nx, ny, ld, mgri, mgrj, nReg,
     &      nRegBrd, nMomBdTp, dBCVal, u, v)
     &      BIND(C, NAME='wd2_vel_bound_cond')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_vel_bound_cond
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nMomBdTp(*)
          REAL(C_DOUBLE) :: dBCVal(*), u(*), v(*)
        END FUNCTION

        FUNCTION wd2_vel_outflow_bcs(nx, ny, ld, mgri, mgrj, nReg,
     &      nRegBrd, nMomBdTp, dBCVal, u, v)
     &      BIND(C, NAME='wd2_vel_outflow_bcs')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_vel_outflow_bcs
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nMomBdTp(*)
          REAL(C_DOUBLE) :: dBCVal(*), u(*), v(*)
        END FUNCTION

        FUNCTION wd2_pres_bound_cond(nx, ny, ld, mgri, mgrj, nReg,
     &      nRegBrd, nRegType, nMomBdTp, dBCVal, p)
     &      BIND(C, NAME='wd2_pres_bound_cond')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_pres_bound_cond
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nRegType(*)
          INTEGER(C_INT) :: nMomBdTp(*)
          REAL(C_DOUBLE) :: dBCVal(*), p(*)
        END FUNCTION

        FUNCTION wd2_dmax_norm(nx, ny, ld, u, res)
     &      BIND(C, NAME='wd2_dmax_norm')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_dmax_norm
          INTEGER(C_INT), VALUE :: nx, ny, ld
          REAL(C_DOUBLE) :: u(*), res
        END FUNCTION

        FUNCTION wd2_diff_max_norm(nx, ny, ld, un, u, res)
     &      BIND(C, NAME='wd2_diff_max_norm')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_diff_max_norm
          INTEGER(C_INT), VALUE :: nx, ny, ld
          REAL(C_DOUBLE) :: un(*), u(*), res
        END FUNCTION

        FUNCTION wd2_temp_bound_cond(nx, ny, ld, mgri, mgrj, nReg,
     &      nRegBrd, nTRgType, nTemBdTp, dTRgVal, dBCVal, t)
     &      BIND(C, NAME='wd2_temp_bound_cond')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_temp_bound_cond
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nTRgType(*)
          INTEGER(C_INT) :: nTemBdTp(*)
          REAL(C_DOUBLE) :: dTRgVal(*), dBCVal(*), t(*)
        END FUNCTION

        FUNCTION wd2_therm_energy(nx, ny, ld, mgri, mgrj, nReg,
     &      nRegBrd, nTRgType, nTemBdTp, dk, pe, dTRgVal, dHGSTval,
     &      dBCVal, rau, rbu, rbv, rgv, djc, xeu, yeu, xzv, yzv,
     &      xec, yec, xzc, yzc, un, vn, u, v, tn, t)
     &      BIND(C, NAME='wd2_therm_energy')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_therm_energy
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nTRgType(*)
          INTEGER(C_INT) :: nTemBdTp(*)
          REAL(C_DOUBLE), VALUE :: dk, pe
          REAL(C_DOUBLE) :: dTRgVal(*), dHGSTval(*), dBCVal(*)
          REAL(C_DOUBLE) :: rau(*), rbu(*), rbv(*), rgv(*), djc(*)
          REAL(C_DOUBLE) :: xeu(*), yeu(*), xzv(*), yzv(*)
          REAL(C_DOUBLE) :: xec(*), yec(*), xzc(*), yzc(*)
          REAL(C_DOUBLE) :: un(*), vn(*), u(*), v(*), tn(*), t(*)
        END FUNCTION

        FUNCTION wd2_eq_state(nx, ny, ld, uref, densref, tmax, tref,
     &      rconst, p, t, den)
     &      BIND(C, NAME='wd2_eq_state')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_eq_state
          INTEGER(C_INT), VALUE :: nx, ny, ld
          REAL(C_DOUBLE), VALUE :: uref, densref, tmax, tref, rconst
          REAL(C_DOUBLE) :: p(*), t(*), den(*)
        END FUNCTION

        FUNCTION wd2_filter(nx, ny, ld, mgri, mgrj, ncomp, nReg,
     &      nRegBrd, nRegType, nMomBdTp, nTRgType, fp, qu)
     &      BIND(C, NAME='wd2_filter')
          IMPORT :: C_INT, C_DOUBLE
          INTEGER(C_INT) :: wd2_filter
          INTEGER(C_INT), VALUE :: nx, ny, ld, mgri, mgrj, ncomp
          INTEGER(C_INT) :: nReg(*), nRegBrd(*), nRegType(*)
          INTEGER(C_INT) :: nMomBdTp(*), nTRgType(*)
          REAL(C_DOUBLE), VALUE :: fp
          REAL(C_DOUBLE) :: qu(*)
        END FUNCTION
      END INTERFACE

      CONTAINS

C     A non-zero status: print the library's message and stop.
      SUBROUTINE wd2_check(ist, who)
        INTEGER(C_INT), INTENT(IN) :: ist
        CHARACTER(LEN=*), INTENT(IN) :: who
        CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
        INTEGER :: k, n
        IF (ist .EQ. 0) RETURN
        CALL C_F_POINTER(wd2_last_error(), msg, [1024])
        n = 0
        DO k = 1, 1024
          IF (msg(k) .EQ. C_NULL_CHAR) EXIT
          n = k
        END DO
        WRITE(*,*) who, ': ', (msg(k), k = 1, n)
        STOP
      END SUBROUTINE
      END MODULE wd2_iface

C-----------------------------------------------------------------------
      INTEGER FUNCTION nAuxMomentum(nx, ny, mqiter, nReg, nRegBrd,
     &    nRegType, nMomBdTp, dk, re, fr, qtol,
     &    dPRporos, dPRporc1, dPRporc2, dBCVal,
     &    ran, rbn, rgn, rac, rbc, rgc, dju, djv,
     &    xec, yec, xzn, yzn, xen, yen, xzc, yzc,
     &    xeu, yeu, xzu, yzu, xev, yev, xzv, yzv,
     &    d, dn, un, vn, us, vs)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny, mqiter
      INTEGER nReg(*), nRegBrd(*), nRegType(*), nMomBdTp(*)
      REAL*8  dk, re, fr, qtol
      REAL*8  dPRporos(*), dPRporc1(*), dPRporc2(*), dBCVal(*)
      REAL*8  ran(*), rbn(*), rgn(*), rac(*), rbc(*), rgc(*)
      REAL*8  dju(*), djv(*)
      REAL*8  xec(*), yec(*), xzn(*), yzn(*)
      REAL*8  xen(*), yen(*), xzc(*), yzc(*)
      REAL*8  xeu(*), yeu(*), xzu(*), yzu(*)
      REAL*8  xev(*), yev(*), xzv(*), yzv(*)
      REAL*8  d(*), dn(*), un(*), vn(*), us(*), vs(*)
      INTEGER(C_INT) ist, nq
      ist = wd2_naux_momentum(nx, ny, mnx + 1, mgri, mgrj, mqiter,
     &    nReg, nRegBrd, nRegType, nMomBdTp, dk, re, fr, qtol,
     &    dPRporos, dPRporc1, dPRporc2, dBCVal,
     &    ran, rbn, rgn, rac, rbc, rgc, dju, djv,
     &    xec, yec, xzn, yzn, xen, yen, xzc, yzc,
     &    xeu, yeu, xzu, yzu, xev, yev, xzv, yzv,
     &    d, dn, un, vn, us, vs, nq)
      CALL wd2_check(ist, 'nAuxMomentum')
      nAuxMomentum = nq
      END

C-----------------------------------------------------------------------
      SUBROUTINE Ppe(nx, ny, nReg, nRegBrd, nRegType, lCartesGrid,
     &    nPpeSolver, msorit, nSorConv, dk, sortol, sorrel,
     &    rau, rbu, rbv, rgv, xeu, yeu, xzv, yzv, u, v, p)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny, nPpeSolver, msorit, nSorConv
      INTEGER nReg(*), nRegBrd(*), nRegType(*)
      LOGICAL lCartesGrid
      REAL*8  dk, sortol, sorrel
      REAL*8  rau(*), rbu(*), rbv(*), rgv(*)
      REAL*8  xeu(*), yeu(*), xzv(*), yzv(*)
      REAL*8  u(*), v(*), p(*)
      INTEGER(C_INT) ist, lc, nc
      lc = 0
      IF (lCartesGrid) lc = 1
      ist = wd2_ppe(nx, ny, mnx + 1, mgri, mgrj, nReg, nRegBrd,
     &    nRegType, lc, nPpeSolver, msorit, nc, dk, sortol, sorrel,
     &    rau, rbu, rbv, rgv, xeu, yeu, xzv, yzv, u, v, p)
      CALL wd2_check(ist, 'Ppe')
      nSorConv = nc
      END

C-----------------------------------------------------------------------
      SUBROUTINE Project(nx, ny, nReg, nRegBrd, nRegType, nMomBdTp,
     &    dk, dju, djv, yeu, xzv, yzu, xev, p, u, v)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nRegType(*), nMomBdTp(*)
      REAL*8  dk
      REAL*8  dju(*), djv(*), yeu(*), xzv(*), yzu(*), xev(*)
      REAL*8  p(*), u(*), v(*)
      INTEGER(C_INT) ist
      ist = wd2_project(nx, ny, mnx + 1, mgri, mgrj, nReg, nRegBrd,
     &    nRegType, nMomBdTp, dk, dju, djv, yeu, xzv, yzu, xev,
     &    p, u, v)
      CALL wd2_check(ist, 'Project')
      END

C-----------------------------------------------------------------------
      SUBROUTINE VelBoundCond(nx, ny, nReg, nRegBrd, nMomBdTp,
     &    dBCVal, u, v)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nMomBdTp(*)
      REAL*8  dBCVal(*), u(*), v(*)
      INTEGER(C_INT) ist
      ist = wd2_vel_bound_cond(nx, ny, mnx + 1, mgri, mgrj, nReg,
     &    nRegBrd, nMomBdTp, dBCVal, u, v)
      CALL wd2_check(ist, 'VelBoundCond')
      END

C-----------------------------------------------------------------------
      SUBROUTINE VelOutflowBCs(nx, ny, nReg, nRegBrd, nMomBdTp,
     &    dBCVal, u, v)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nMomBdTp(*)
      REAL*8  dBCVal(*), u(*), v(*)
      INTEGER(C_INT) ist
      ist = wd2_vel_outflow_bcs(nx, ny, mnx + 1, mgri, mgrj, nReg,
     &    nRegBrd, nMomBdTp, dBCVal, u, v)
      CALL wd2_check(ist, 'VelOutflowBCs')
      END

C-----------------------------------------------------------------------
      SUBROUTINE PresBoundCond(nx, ny, nReg, nRegBrd, nRegType,
     &    nMomBdTp, dBCVal, p)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nRegType(*), nMomBdTp(*)
      REAL*8  dBCVal(*), p(*)
      INTEGER(C_INT) ist
      ist = wd2_pres_bound_cond(nx, ny, mnx + 1, mgri, mgrj, nReg,
     &    nRegBrd, nRegType, nMomBdTp, dBCVal, p)
      CALL wd2_check(ist, 'PresBoundCond')
      END

C-----------------------------------------------------------------------
      REAL*8 FUNCTION DMaxNorm(nx, ny, u)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      REAL*8  u(*)
      INTEGER(C_INT) ist
      REAL(C_DOUBLE) r
      ist = wd2_dmax_norm(nx, ny, mnx + 1, u, r)
      CALL wd2_check(ist, 'DMaxNorm')
      DMaxNorm = r
      END

C-----------------------------------------------------------------------
      REAL*8 FUNCTION DiffMaxNorm(nx, ny, un, u)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      REAL*8  un(*), u(*)
      INTEGER(C_INT) ist
      REAL(C_DOUBLE) r
      ist = wd2_diff_max_norm(nx, ny, mnx + 1, un, u, r)
      CALL wd2_check(ist, 'DiffMaxNorm')
      DiffMaxNorm = r
      END

C-----------------------------------------------------------------------
      SUBROUTINE TempBoundCond(nx, ny, nReg, nRegBrd, nTRgType,
     &    nTemBdTp, dTRgVal, dBCVal, t)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nTRgType(*), nTemBdTp(*)
      REAL*8  dTRgVal(*), dBCVal(*), t(*)
      INTEGER(C_INT) ist
      ist = wd2_temp_bound_cond(nx, ny, mnx + 1, mgri, mgrj, nReg,
     &    nRegBrd, nTRgType, nTemBdTp, dTRgVal, dBCVal, t)
      CALL wd2_check(ist, 'TempBoundCond')
      END

C-----------------------------------------------------------------------
      SUBROUTINE ThermEnergy(nx, ny, nReg, nRegBrd, nTRgType,
     &    nTemBdTp, dk, pe, dTRgVal, dHGSTval, dBCVal,
     &    rau, rbu, rbv, rgv, djc, xeu, yeu, xzv, yzv,
     &    xec, yec, xzc, yzc, un, vn, u, v, tn, t)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      INTEGER nReg(*), nRegBrd(*), nTRgType(*), nTemBdTp(*)
      REAL*8  dk, pe
      REAL*8  dTRgVal(*), dHGSTval(*), dBCVal(*)
      REAL*8  rau(*), rbu(*), rbv(*), rgv(*), djc(*)
      REAL*8  xeu(*), yeu(*), xzv(*), yzv(*)
      REAL*8  xec(*), yec(*), xzc(*), yzc(*)
      REAL*8  un(*), vn(*), u(*), v(*), tn(*), t(*)
      INTEGER(C_INT) ist
      ist = wd2_therm_energy(nx, ny, mnx + 1, mgri, mgrj, nReg,
     &    nRegBrd, nTRgType, nTemBdTp, dk, pe, dTRgVal, dHGSTval,
     &    dBCVal, rau, rbu, rbv, rgv, djc, xeu, yeu, xzv, yzv,
     &    xec, yec, xzc, yzc, un, vn, u, v, tn, t)
      CALL wd2_check(ist, 'ThermEnergy')
      END

C-----------------------------------------------------------------------
      SUBROUTINE EqState(nx, ny, uref, densref, tmax, tref, rconst,
     &    p, t, den)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny
      REAL*8  uref, densref, tmax, tref, rconst
      REAL*8  p(*), t(*), den(*)
      INTEGER(C_INT) ist
      ist = wd2_eq_state(nx, ny, mnx + 1, uref, densref, tmax, tref,
     &    rconst, p, t, den)
      CALL wd2_check(ist, 'EqState')
      END

C-----------------------------------------------------------------------
      SUBROUTINE Filter(nx, ny, ncomp, nReg, nRegBrd, nRegType,
     &    nMomBdTp, nTRgType, fp, qu)
      USE wd2_iface
      IMPLICIT NONE
      INCLUDE 'config.f'
      INTEGER nx, ny, ncomp
      INTEGER nReg(*), nRegBrd(*), nRegType(*), nMomBdTp(*)
      INTEGER nTRgType(*)
      REAL*8  fp, qu(*)
      INTEGER(C_INT) ist
      ist = wd2_filter(nx, ny, mnx + 1, mgri, mgrj, ncomp, nReg,
     &    nRegBrd, nRegType, nMomBdTp, nTRgType, fp, qu)
      CALL wd2_check(ist, 'Filter')
      END
