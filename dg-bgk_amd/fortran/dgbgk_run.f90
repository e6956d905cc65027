! dgbgk_run.f90 -- Fortran 90 host for the DG-BGK time march on MI355X.
!
! Usage:  dgbgk_run <deck.con> [--fp32] [--jacobi] [--lookup=reference|corrected]
!                              [--check] [--restart-out]
!
! Reads the reference's run deck (namelist /Input/ IVD, the same keys and
! defaults as src/spectralbgk.f90:116-148), its velocity file
! (src/spectralbgk.f90:157-161, src/vspace.f90:16-18) and its mesh file
! (src/spectralbgk.f90:240-272, src/gtinpt.f90:24-50); writes FILE.OUT,
! RESIDUAL.RES and RESULTS1/RESULTS2.RES in the reference's layouts
! (src/advnce.f90:103-120,163-170,403-435,468-489).  The time march itself
! runs on the GPU through the C ABI (module DGBGK_API).  --check builds the
! host-side setup only and prints its sizes (no GPU needed).
!
! Restart: with IVD%RS = 1 the initial f is read from IVD%RestartInFile;
! --restart-out writes f to IVD%RestartOutFile and the mesh to
! IVD%GIDMeshFile after the loop.  The layout is the reference writer's
! (src/output.f90:54-58,110-125: a 'NPOIN     VNPNT' line, the two counts,
! then per ordinate and element "IE CX CY f1 f2 f3") at full precision,
! and the reader takes the last three numbers of each line.  (The
! reference's writer is not called and its reader, src/inicon2.f90:87-110,
! reads the first three numbers of such a line as f, so the two do not
! round-trip.)
PROGRAM DGBGK_RUN
  USE ISO_C_BINDING
  USE DGBGK_API
  IMPLICIT NONE

  TYPE InputVariables
    INTEGER :: TORDER, NTIME, FORCEOUT, IMMAT, RS, INF, NVSPACEPART
    REAL(8) :: CSAFM, rv, T1, P1, U0, V0, CINF(4), W, ALPHA, R, d, M
    CHARACTER(len=80) :: LobattoFile, PSpaceFile, OutFile, PartitionFile
    CHARACTER(len=80) :: RestartInFile, ResidualFile, ResultsFile1, ResultsFile2
    CHARACTER(len=80) :: RestartOutFile, GIDMeshFile
  END TYPE
  TYPE(InputVariables) :: IVD
  NAMELIST /Input/ IVD

  CHARACTER(len=256) :: DECK, ARG
  CHARACTER(len=80) :: TEXT
  INTEGER :: NARG, IA, I, J, IDUM, NLINES, NDIMN, NELEM, NPOIN, NBOUN, RORDER, TORDER
  INTEGER :: ITIME, NTIME, NB, IP, IOS, IST, K
  INTEGER(C_INT) :: ST, NSIDE, NVEL, NT, NW, RS, IT, NE, NP, NBC
  INTEGER(C_INT), ALLOCATABLE :: INTMA(:, :), BSIDO(:, :)
  REAL(C_DOUBLE), ALLOCATABLE :: COORD(:, :), ETA(:), WT(:)
  REAL(C_DOUBLE), ALLOCATABLE :: ND(:), RHO(:), UVEL(:), VVEL(:), PS(:), TEMP(:)
  TYPE(DGBGK_PARAMS_T) :: PRM
  TYPE(DGBGK_STEP_INFO_T), ALLOCATABLE :: INFO(:)
  TYPE(C_PTR) :: CTX, SETUP
  LOGICAL :: CHECK, RSOUT
  REAL(8) :: T0, T1S, SUMW
  REAL(C_DOUBLE), ALLOCATABLE :: F(:, :, :)

  PRM%PRECISION = DGBGK_FP64
  PRM%ORDER = DGBGK_ORDER_REFERENCE
  PRM%LOOKUP = DGBGK_LOOKUP_CORRECTED
  PRM%VCHUNK = 0
  CHECK = .FALSE.
  RSOUT = .FALSE.
  NARG = COMMAND_ARGUMENT_COUNT()
  IF (NARG .LT. 1) THEN
    WRITE (*, *) 'usage: dgbgk_run <deck.con> [--fp32] [--jacobi] [--lookup=reference]'
    STOP 2
  ENDIF
  CALL GET_COMMAND_ARGUMENT(1, DECK)
  DO IA = 2, NARG
    CALL GET_COMMAND_ARGUMENT(IA, ARG)
    SELECT CASE (TRIM(ARG))
    CASE ('--fp32')
      PRM%PRECISION = DGBGK_FP32
    CASE ('--jacobi')
      PRM%ORDER = DGBGK_ORDER_JACOBI
    CASE ('--lookup=reference')
      PRM%LOOKUP = DGBGK_LOOKUP_REFERENCE
    CASE ('--lookup=corrected')
      PRM%LOOKUP = DGBGK_LOOKUP_CORRECTED
    CASE ('--check')
      CHECK = .TRUE.
    CASE ('--restart-out')
      RSOUT = .TRUE.
    CASE DEFAULT
      WRITE (*, *) 'unknown option ', TRIM(ARG)
      STOP 2
    END SELECT
  ENDDO

  ! deck defaults (src/spectralbgk.f90:116-147)
  IVD%TORDER = 20; IVD%NTIME = 10000; IVD%FORCEOUT = 0; IVD%IMMAT = 1
  IVD%INF = 1; IVD%RS = 0; IVD%CSAFM = 0.5D0; IVD%rv = 2000D0; IVD%T1 = 293D0
  IVD%P1 = 0.1D0; IVD%U0 = 100D0; IVD%V0 = 0D0; IVD%W = 1D0; IVD%ALPHA = 0.9D0
  IVD%R = 287D0; IVD%d = 250D-12; IVD%M = 32D0
  IVD%CINF = (/294D0, 0.001473D0, 171.85D0, 0D0/)
  IVD%NVSPACEPART = 1
  IVD%LobattoFile = 'Lobatto20.txt'; IVD%PSpaceFile = 'AEROFOIL.RES'
  IVD%OutFile = 'FILE.OUT'; IVD%PartitionFile = 'METIS3WAYPARTAEROFOIL13.RES'
  IVD%RestartInFile = 'RESTART.RES'; IVD%ResidualFile = 'RESIDUAL.RES'
  IVD%ResultsFile1 = 'RESULTS1.RES'; IVD%ResultsFile2 = 'RESULTS2.RES'
  IVD%RestartOutFile = 'RESTART.RES'; IVD%GIDMeshFile = 'GIDMESH.RES'
  PRINT *, 'CONFIGURATION FILE = ', TRIM(DECK)
  OPEN (13, FILE=DECK, STATUS='OLD')
  READ (13, Input)
  CLOSE (13)

  ! velocity file: RORDER, then RORDER lines "eta weight"
  OPEN (10, FILE=IVD%LobattoFile, STATUS='OLD')
  READ (10, *) RORDER
  ALLOCATE (ETA(RORDER), WT(RORDER))
  DO I = 1, RORDER
    READ (10, *) ETA(I), WT(I)
  ENDDO
  CLOSE (10)
  TORDER = IVD%TORDER
  SUMW = 0D0
  DO I = 1, RORDER
    DO J = 1, TORDER
      SUMW = SUMW + WT(I)*1.0D0
    ENDDO
  ENDDO
  PRINT *, 'SUMWEIGHT=', SUMW

  ! mesh file
  OPEN (14, FILE=IVD%PSpaceFile, STATUS='OLD')
  OPEN (15, FILE=IVD%OutFile, STATUS='UNKNOWN')
  READ (14, *) NLINES
  DO I = 1, NLINES
    READ (14, *) TEXT
    WRITE (15, 1) TEXT
  ENDDO
  READ (14, *) TEXT
  WRITE (15, 1) TEXT
  READ (14, *) NDIMN
  WRITE (15, 5) NDIMN
  READ (14, *) TEXT
  WRITE (15, 1) TEXT
  READ (14, *) NELEM, NPOIN, NBOUN
  WRITE (15, 5) NELEM, NPOIN, NBOUN
  WRITE (15, 170)
  WRITE (15, 180)
  WRITE (15, 5) RORDER, TORDER, RORDER*TORDER
  CLOSE (15)
  ALLOCATE (INTMA(3, NELEM), COORD(2, NPOIN), BSIDO(6, NBOUN))
  READ (14, '(A)') TEXT
  DO I = 1, NELEM
    READ (14, *) IDUM, (INTMA(J, I), J=1, 3)
  ENDDO
  READ (14, '(A)') TEXT
  DO I = 1, NPOIN
    READ (14, *) IDUM, (COORD(J, I), J=1, 2)
  ENDDO
  READ (14, '(A)') TEXT
  DO I = 1, NBOUN
    READ (14, *) (BSIDO(J, I), J=1, 6)
  ENDDO
  CLOSE (14)

  PRM%RV = IVD%rv; PRM%CSAFM = IVD%CSAFM; PRM%ALPHA = IVD%ALPHA; PRM%RGAS = IVD%R
  PRM%D = IVD%d; PRM%M = IVD%M; PRM%T1 = IVD%T1; PRM%P1 = IVD%P1
  PRM%U0 = IVD%U0; PRM%V0 = IVD%V0; PRM%CINF = IVD%CINF; PRM%INF = IVD%INF
  PRM%IMMAT = IVD%IMMAT

  IF (CHECK) THEN
    ST = DGBGK_SETUP_CREATE(NELEM, NPOIN, NBOUN, INTMA, COORD, BSIDO, RORDER, TORDER, &
                            ETA, WT, PRM, SETUP)
    IF (ST .NE. 0) THEN
      WRITE (*, *) 'SETUP FAILED ', ST, ' ', DGBGK_LAST_ERROR()
      STOP 1
    ENDIF
    ST = DGBGK_SETUP_SIZES(SETUP, NSIDE, NVEL, NT, NW, RS)
    WRITE (*, '(A,7I10)') 'CHECK', NELEM, NPOIN, NBOUN, NSIDE, NVEL, NT, NW
    CALL DGBGK_SETUP_DESTROY(SETUP)
    STOP
  ENDIF

  ST = DGBGK_CREATE(NELEM, NPOIN, NBOUN, INTMA, COORD, BSIDO, RORDER, TORDER, ETA, WT, &
                    PRM, CTX)
  IF (ST .NE. 0) THEN
    WRITE (*, *) 'DGBGK_CREATE FAILED ', ST, ' ', DGBGK_LAST_ERROR()
    STOP 1
  ENDIF
  ALLOCATE (ND(NPOIN), RHO(NPOIN), UVEL(NPOIN), VVEL(NPOIN), PS(NPOIN), TEMP(NPOIN))
  IF (IVD%RS .EQ. 1) THEN
    ! restart input replaces the INICON2 state (src/inicon2.f90:61-110)
    CALL READ_RESTART()
    ST = DGBGK_SET_F(CTX, F)
    CALL CHECKST(ST)
  ENDIF

  PRINT *, 'OPENING RESIDUAL FILE = ', TRIM(IVD%ResidualFile)
  OPEN (16, FILE=IVD%ResidualFile, STATUS='UNKNOWN')
  OPEN (17, FILE=IVD%ResultsFile1, STATUS='UNKNOWN')
  OPEN (18, FILE=IVD%ResultsFile2, STATUS='UNKNOWN')
  ! step-0 fields: initial MACROS + GETMAC (src/advnce.f90:139-169)
  ST = DGBGK_POINT_FIELDS(CTX, ND, RHO, UVEL, VVEL, PS, TEMP)
  CALL CHECKST(ST)
  CALL WRITE_RESULTS(0)
  NTIME = IVD%NTIME
  PRINT *, 'NTIME= ', NTIME
  ALLOCATE (INFO(1000))
  ITIME = 0
  CALL CPU_TIME(T0)
  DO WHILE (ITIME .LT. NTIME)
    ! advance to the next gather step (every 1000 steps) or the end
    NB = MIN(NTIME - ITIME, 1000 - MOD(ITIME, 1000))
    ST = DGBGK_STEP(CTX, NB, INFO)
    CALL CHECKST(ST)
    DO K = 1, NB
      IT = ITIME + K
      IF (IVD%FORCEOUT .EQ. 0) THEN
        WRITE (16, *) IT, INFO(K)%LOGRES
      ELSE
        WRITE (16, *) IT, INFO(K)%LOGRES, INFO(K)%LIFT, INFO(K)%DRAG
      ENDIF
    ENDDO
    ITIME = ITIME + NB
    IF (MOD(ITIME, 1000) .EQ. 0) THEN
      ST = DGBGK_POINT_FIELDS(CTX, ND, RHO, UVEL, VVEL, PS, TEMP)
      CALL CHECKST(ST)
      CALL WRITE_RESULTS(ITIME)
    ENDIF
    CALL CPU_TIME(T1S)
    WRITE (*, '(A,I8,A,ES12.4,A,F10.2)') 'TIMESTEP = ', ITIME, '  SUMRES = ', &
      INFO(NB)%SUMRES, '  host s = ', T1S - T0
  ENDDO
  WRITE (*, *) 'END OF TIME LOOP'
  ! after the loop the reference writes the last gathered fields with
  ! ITIME = NTIME+1 (src/advnce.f90:468-476)
  CALL WRITE_RESULTS(NTIME + 1)
  CLOSE (16)
  CLOSE (17)
  CLOSE (18)
  IF (RSOUT) THEN
    CALL WRITE_RESTART()
    CALL WRITE_GIDMESH()
  ENDIF
  IST = DGBGK_SIZES(CTX, NSIDE, NVEL, NE, NP, NBC, IT)
  CALL DGBGK_DESTROY(CTX)

1 FORMAT(20A30)
5 FORMAT(I6, I6, I6)
170 FORMAT('V-SPACE STATS:')
180 FORMAT('LOBATTO ORDER      TOTAL NUMBER OF NODES')

CONTAINS

  SUBROUTINE CHECKST(S)
    INTEGER(C_INT), INTENT(IN) :: S
    IF (S .NE. 0) THEN
      WRITE (*, *) 'DGBGK ERROR ', S, ' ', DGBGK_LAST_ERROR()
      STOP 1
    ENDIF
  END SUBROUTINE

  SUBROUTINE READ_RESTART()
    INTEGER :: T1N, T2N, IV, IE, IU
    REAL(8) :: CXR, CYR
    IF (.NOT. ALLOCATED(F)) ALLOCATE (F(3, RORDER*TORDER, NELEM))
    PRINT *, 'READING RESTART FILE = ', TRIM(IVD%RestartInFile)
    OPEN (11, FILE=IVD%RestartInFile, STATUS='OLD')
    READ (11, *) TEXT
    READ (11, *) T1N, T2N
    IF ((T1N .NE. NPOIN) .OR. (T2N .NE. RORDER*TORDER)) THEN
      WRITE (*, *) 'RESTART FILE DOES NOT MATCH: NPOIN, VNPNT =', T1N, T2N
      STOP 1
    ENDIF
    DO IV = 1, RORDER*TORDER
      DO IE = 1, NELEM
        READ (11, *) IU, CXR, CYR, F(1, IV, IE), F(2, IV, IE), F(3, IV, IE)
      ENDDO
    ENDDO
    CLOSE (11)
  END SUBROUTINE

  SUBROUTINE WRITE_RESTART()
    INTEGER :: IV, IE, IR, IZ
    REAL(8) :: RT, TH
    IF (.NOT. ALLOCATED(F)) ALLOCATE (F(3, RORDER*TORDER, NELEM))
    ST = DGBGK_GET_F(CTX, F)
    CALL CHECKST(ST)
    PRINT *, 'OPENING NEW RESTART FILE = ', TRIM(IVD%RestartOutFile)
    OPEN (20, FILE=IVD%RestartOutFile, STATUS='UNKNOWN')
    WRITE (20, '(A)') 'NPOIN     VNPNT'
    WRITE (20, *) NPOIN, '    ', RORDER*TORDER
    DO IV = 1, RORDER*TORDER
      ! ordinate IV = (IR-1)*TORDER + IZ (src/vspace.f90:39-49), mapped to
      ! polar and Cartesian coordinates as src/output.f90:112-118
      IR = (IV - 1)/TORDER + 1
      IZ = IV - (IR - 1)*TORDER
      RT = ETA(IR)*(IVD%rv/2) + (IVD%rv/2)
      TH = (-1D0 + (2D0*IZ)/TORDER)*3.1416D0
      DO IE = 1, NELEM
        WRITE (20, 700) IE, RT*COS(TH), RT*SIN(TH), F(1, IV, IE), F(2, IV, IE), F(3, IV, IE)
      ENDDO
    ENDDO
    CLOSE (20)
700 FORMAT(I10, 2(1X, ES15.7), 3(1X, ES25.17E3))
  END SUBROUTINE

  SUBROUTINE WRITE_GIDMESH()
    ! src/output.f90:130-153 and its formats 200-240
    INTEGER :: Q
    OPEN (19, FILE=IVD%GIDMeshFile, STATUS='UNKNOWN')
    WRITE (19, '(A)') 'mesh dimension = 2 elemtype triangle nnode = 3'
    WRITE (19, '(A)') '   COORDINATES'
    DO Q = 1, NPOIN
      WRITE (19, '(I10, 2(2X, F10.5))') Q, COORD(1, Q), COORD(2, Q)
    ENDDO
    WRITE (19, '(A)') 'end coordinates'
    WRITE (19, '(A)') 'elements'
    DO Q = 1, NELEM
      WRITE (19, '(10I8)') Q, INTMA(1, Q), INTMA(2, Q), INTMA(3, Q)
    ENDDO
    WRITE (19, '(A)') 'end elements'
    CLOSE (19)
  END SUBROUTINE

  SUBROUTINE WRITE_RESULTS(IT)
    INTEGER, INTENT(IN) :: IT
    INTEGER :: Q
    WRITE (17, 500) IT
    WRITE (18, 500) IT
    DO Q = 1, NPOIN
      WRITE (17, 600) Q, ND(Q), UVEL(Q), VVEL(Q)
      WRITE (18, 600) Q, RHO(Q), PS(Q), TEMP(Q)
    ENDDO
500 FORMAT('RESULTS        1   ', I5, '    2      1      0')
600 FORMAT(I5, 4X, 3(1pe12.5, 1X))
  END SUBROUTINE

END PROGRAM DGBGK_RUN
