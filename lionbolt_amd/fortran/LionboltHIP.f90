!  LionboltHIP: iso_c_binding shim over liblbhip (include/lionbolt.h).
!
!  The reference's operators (src/operators/TransportOperator.f90,
!  ScatteringOperator.f90, BoltzmannOperator.f90) keep their type-bound
!  interface Build / MatVec / MatInv / Destroy; their bodies call the procedures
!  below instead of Sweep / SNMGXS / GMRESWrapper.  HIPOperator holds one
!  library handle; the Build* procedures take the MeshClass / AngularClass /
!  XSType fields they need as plain arrays (1-based Fortran indices, boundary
!  neighbour 0, as the reference stores them).  Every failure ends in
!  stophere-like termination with the library's message, keeping the
!  reference's abort semantics (src/solvers/SPARSKIT.f90:159-169).
Module LionboltHIP
    use iso_c_binding
    Implicit None
    Private

    Integer (c_int), Parameter, Public :: LB_SOLVER_GMRES = 0, LB_SOLVER_SI = 1
    Integer (c_int), Parameter, Public :: LB_SPACE_PSI = 0, LB_SPACE_MOMENT = 1

    Type, Bind(C), Public :: lb_mesh_desc
        Integer (c_int64_t) :: NE, NK
        Type (c_ptr)        :: rg, node, nbr, mat
        Integer (c_int)     :: index_base
    End Type

    Type, Bind(C), Public :: lb_quad_desc
        Integer (c_int64_t) :: NI
        Type (c_ptr)        :: k, w
        Integer (c_int)     :: L
    End Type

    Type, Bind(C), Public :: lb_solver_opts
        Integer (c_int) :: kind, space, restart, maxit
        Real (c_double) :: rtol, atol
        Integer (c_int) :: verbose
    End Type

    Integer (c_int), Parameter, Public :: LB_CUTOUT_NONE = 0, LB_CUTOUT_CIRCLE = 1, LB_CUTOUT_RECTANGLE = 2

    ! FieldGeometry + ExternalBeam of a planar beam (src/sources/Sources.f90:13-80):
    ! unit axis, field origin, the field plane's axes e1, e2 (the z-y-z rotation of
    ! the axis for a rectangle), the cutout and its parameters, the population n.
    Type, Bind(C), Public :: lb_beam_desc
        Real (c_double) :: axis (3), origin (3), e1 (3), e2 (3)
        Integer (c_int) :: cutout
        Real (c_double) :: params (4)
        Real (c_double) :: n
    End Type

    Type, Bind(C), Public :: lb_solve_stats
        Integer (c_int) :: status, iterates, matvecs, sweeps
        Real (c_double) :: final_norm, true_residual, seconds
        Integer (c_int) :: fallback
    End Type

    Interface
        Integer (c_int) Function lb_op_create (mesh, quad, op) Bind(C, name='lb_op_create')
            import :: c_int, c_ptr, lb_mesh_desc, lb_quad_desc
            Type (lb_mesh_desc), Intent (In)  :: mesh
            Type (lb_quad_desc), Intent (In)  :: quad
            Type (c_ptr),        Intent (Out) :: op
        End Function
        Integer (c_int) Function lb_op_set_physics (op, Nm, sigt, L, sigs) Bind(C, name='lb_op_set_physics')
            import :: c_int, c_ptr, c_double
            Type (c_ptr),    Value       :: op
            Integer (c_int), Value       :: Nm, L
            Real (c_double), Intent (In) :: sigt (*), sigs (*)
        End Function
        Integer (c_int) Function lb_op_build_delta_down (op, Nm, xsdd) Bind(C, name='lb_op_build_delta_down')
            import :: c_int, c_ptr, c_double
            Type (c_ptr),    Value       :: op
            Integer (c_int), Value       :: Nm
            Real (c_double), Intent (In) :: xsdd (*)
        End Function
        Integer (c_int) Function lb_transport_matinv (op, s, n) Bind(C, name='lb_transport_matinv')
            import :: c_int, c_ptr, c_double, c_int64_t
            Type (c_ptr),        Value          :: op
            Real (c_double),     Intent (InOut) :: s (*)
            Integer (c_int64_t), Value          :: n
        End Function
        Integer (c_int) Function lb_scatter_matvec (op, x, n) Bind(C, name='lb_scatter_matvec')
            import :: c_int, c_ptr, c_double, c_int64_t
            Type (c_ptr),        Value          :: op
            Real (c_double),     Intent (InOut) :: x (*)
            Integer (c_int64_t), Value          :: n
        End Function
        Integer (c_int) Function lb_boltzmann_matvec (op, x, n) Bind(C, name='lb_boltzmann_matvec')
            import :: c_int, c_ptr, c_double, c_int64_t
            Type (c_ptr),        Value          :: op
            Real (c_double),     Intent (InOut) :: x (*)
            Integer (c_int64_t), Value          :: n
        End Function
        Integer (c_int) Function lb_boltzmann_matinv (op, s, n, opts, stats) Bind(C, name='lb_boltzmann_matinv')
            import :: c_int, c_ptr, c_double, c_int64_t, lb_solver_opts, lb_solve_stats
            Type (c_ptr),          Value          :: op
            Real (c_double),       Intent (InOut) :: s (*)
            Integer (c_int64_t),   Value          :: n
            Type (lb_solver_opts), Intent (In)    :: opts
            Type (lb_solve_stats), Intent (Out)   :: stats
        End Function
        Integer (c_int) Function lb_fluence (op, v, phi) Bind(C, name='lb_fluence')
            import :: c_int, c_ptr, c_double
            Type (c_ptr),    Value        :: op
            Real (c_double), Intent (In)  :: v (*)
            Real (c_double), Intent (Out) :: phi (*)
        End Function
        Integer (c_int) Function lb_beam_source (op, beam, s, n) Bind(C, name='lb_beam_source')
            import :: c_int, c_ptr, c_double, c_int64_t, lb_beam_desc
            Type (c_ptr),        Value          :: op
            Type (lb_beam_desc), Intent (In)    :: beam
            Real (c_double),     Intent (InOut) :: s (*)
            Integer (c_int64_t), Value          :: n
        End Function
        Integer (c_int) Function lb_beam_add_angular_fluence (op, beam, attn, unc, n) &
                Bind(C, name='lb_beam_add_angular_fluence')
            import :: c_int, c_ptr, c_double, c_int64_t, lb_beam_desc
            Type (c_ptr),        Value          :: op
            Type (lb_beam_desc), Intent (In)    :: beam
            Real (c_double),     Intent (In)    :: attn (*)
            Real (c_double),     Intent (InOut) :: unc (*)
            Integer (c_int64_t), Value          :: n
        End Function
        Subroutine lb_op_destroy (op) Bind(C, name='lb_op_destroy')
            import :: c_ptr
            Type (c_ptr), Value :: op
        End Subroutine
        Type (c_ptr) Function lb_last_error () Bind(C, name='lb_last_error')
            import :: c_ptr
        End Function
    End Interface

    Type, Public :: HIPOperator
        Type (c_ptr)                   :: h = c_null_ptr
        Integer (c_int64_t)            :: n = 0      ! NENK * NI
        Integer (c_int64_t)            :: NENK = 0
        Logical                        :: PHYSREADY = .FALSE.
        Character (LEN=:), Allocatable :: solver     ! 'GMRES' or 'SI' (AngularClass%solver)
        Integer (c_int)                :: space = LB_SPACE_PSI
        Integer (c_int)                :: restart = 10
        Type (lb_solve_stats)          :: stats
    Contains
        Procedure :: BuildSpaceAngle
        Procedure :: BuildPhysics
        Procedure :: BuildDeltaDown
        Procedure :: TransportMatInv
        Procedure :: ScatterMatVec
        Procedure :: BoltzmannMatVec
        Procedure :: BoltzmannMatInv
        Procedure :: Fluence
        Procedure :: BeamSource
        Procedure :: BeamAddAngularFluence
        Procedure :: Destroy
    End Type

    Public :: lb_stop

Contains

    Subroutine lb_stop (who, rc)
        Character (*),   Intent (In) :: who
        Integer (c_int), Intent (In) :: rc
        Character (kind=c_char), Pointer :: msg (:)
        Character (LEN=512) :: text
        Integer :: i
        text = ''
        call c_f_pointer(lb_last_error(), msg, [512])
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        write (*,'(A,A,I0,A,A)') who, ': error ', rc, ': ', TRIM(text)
        error stop 1
    End Subroutine

    ! Op%Build(mesh) + Op%Build(mesh, angular): rg(3,NK), node(4,NE), neighbors(4,NE), mat(NE), k(3,NI), w(NI), L
    Subroutine BuildSpaceAngle (Op, rg, node, neighbors, mat, k, w, L, solver)
        Class (HIPOperator),         Intent (InOut) :: Op
        Real (c_double),     Target, Intent (In)    :: rg (:,:), k (:,:), w (:)
        Integer (c_int64_t), Target, Intent (In)    :: node (:,:), neighbors (:,:), mat (:)
        Integer,                     Intent (In)    :: L
        Character (*),               Intent (In)    :: solver
        Type (lb_mesh_desc) :: md
        Type (lb_quad_desc) :: qd
        Integer (c_int)     :: rc
        md%NE = SIZE(node, DIM=2)
        md%NK = SIZE(rg, DIM=2)
        md%rg = c_loc(rg)
        md%node = c_loc(node)
        md%nbr = c_loc(neighbors)
        md%mat = c_loc(mat)
        md%index_base = 1
        qd%NI = SIZE(k, DIM=2)
        qd%k = c_loc(k)
        qd%w = c_loc(w)
        qd%L = L
        call Op%Destroy()
        rc = lb_op_create(md, qd, Op%h)
        if (rc /= 0) call lb_stop('Op%Build(mesh, angular)', rc)
        Op%NENK = 4 * md%NE
        Op%n = Op%NENK * qd%NI
        Op%solver = solver
    End Subroutine

    ! Op%Build(angular, XS): XS%t(Nm), XS%s(0:L, Nm)
    Subroutine BuildPhysics (Op, sigt, sigs)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (In)    :: sigt (:), sigs (0:,:)
        Integer (c_int) :: rc
        rc = lb_op_set_physics(Op%h, INT(SIZE(sigt), c_int), sigt, INT(SIZE(sigs, DIM=1) - 1, c_int), sigs)
        if (rc /= 0) call lb_stop('Op%Build(angular, XS)', rc)
        Op%PHYSREADY = .TRUE.
    End Subroutine

    ! ScatteringOp%BuildDeltaDown(XSdd): K = XSdd(mat) I for the next MatVec calls
    Subroutine BuildDeltaDown (Op, XSdd)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (In)    :: XSdd (:)
        Integer (c_int) :: rc
        rc = lb_op_build_delta_down(Op%h, INT(SIZE(XSdd), c_int), XSdd)
        if (rc /= 0) call lb_stop('ScatteringOp%BuildDeltaDown', rc)
        Op%PHYSREADY = .TRUE.
    End Subroutine

    ! TransportOp%MatInv: s <- T^-1 s
    Subroutine TransportMatInv (Op, s)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (InOut) :: s (:)
        Integer (c_int) :: rc
        rc = lb_transport_matinv(Op%h, s, INT(SIZE(s), c_int64_t))
        if (rc /= 0) call lb_stop('TransportOp%MatInv', rc)
    End Subroutine

    ! ScatteringOp%MatVec: x <- K x
    Subroutine ScatterMatVec (Op, x)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (InOut) :: x (:)
        Integer (c_int) :: rc
        rc = lb_scatter_matvec(Op%h, x, INT(SIZE(x), c_int64_t))
        if (rc /= 0) call lb_stop('ScatteringOp%MatVec', rc)
    End Subroutine

    ! BoltzmannOp%MatVec: x <- x - T^-1 K x
    Subroutine BoltzmannMatVec (Op, x)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (InOut) :: x (:)
        Integer (c_int) :: rc
        rc = lb_boltzmann_matvec(Op%h, x, INT(SIZE(x), c_int64_t))
        if (rc /= 0) call lb_stop('BoltzmannOp%MatVec', rc)
    End Subroutine

    ! BoltzmannOp%MatInv: s (source) <- psi, GMRES(10) or SI as the reference
    Subroutine BoltzmannMatInv (Op, s)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (InOut) :: s (:)
        Type (lb_solver_opts) :: o
        Integer (c_int) :: rc
        o%kind = MERGE(LB_SOLVER_SI, LB_SOLVER_GMRES, Op%solver == 'SI')
        o%space = Op%space
        o%restart = Op%restart
        o%maxit = -1
        o%rtol = 1.0e-8_c_double
        o%atol = 1.0e-12_c_double
        o%verbose = 1
        rc = lb_boltzmann_matinv(Op%h, s, INT(SIZE(s), c_int64_t), o, Op%stats)
        if (rc /= 0) call lb_stop('BoltzmannOp%MatInv', rc)
    End Subroutine

    ! SpaceAngleVector%Fluence
    Function Fluence (Op, v) Result (phi)
        Class (HIPOperator), Intent (InOut) :: Op
        Real (c_double),     Intent (In)    :: v (:)
        Real (c_double), Allocatable        :: phi (:)
        Integer (c_int) :: rc
        ALLOCATE(phi(Op%NENK))
        rc = lb_fluence(Op%h, v, phi)
        if (rc /= 0) call lb_stop('SpaceAngleVector%Fluence', rc)
    End Function

    ! ExternalBeam%Source => FCSSource (planar): s <- s + the first-collision source,
    ! attenuated by the operator's sigma_t and weighted by its sigma_l, as the
    ! reference does with OpK
    Subroutine BeamSource (Op, beam, s)
        Class (HIPOperator), Intent (InOut) :: Op
        Type (lb_beam_desc), Intent (In)    :: beam
        Real (c_double),     Intent (InOut) :: s (:)
        Integer (c_int) :: rc
        rc = lb_beam_source(Op%h, beam, s, INT(SIZE(s), c_int64_t))
        if (rc /= 0) call lb_stop('ExternalBeam%Source', rc)
    End Subroutine

    ! ExternalBeam%AddAngularFluence (planar): unc <- unc + the uncollided angular fluence
    Subroutine BeamAddAngularFluence (Op, beam, attn, unc)
        Class (HIPOperator), Intent (InOut) :: Op
        Type (lb_beam_desc), Intent (In)    :: beam
        Real (c_double),     Intent (In)    :: attn (:)
        Real (c_double),     Intent (InOut) :: unc (:)
        Integer (c_int) :: rc
        rc = lb_beam_add_angular_fluence(Op%h, beam, attn, unc, INT(SIZE(unc), c_int64_t))
        if (rc /= 0) call lb_stop('ExternalBeam%AddAngularFluence', rc)
    End Subroutine

    Subroutine Destroy (Op)
        Class (HIPOperator), Intent (InOut) :: Op
        if (c_associated(Op%h)) call lb_op_destroy(Op%h)
        Op%h = c_null_ptr
        Op%PHYSREADY = .FALSE.
    End Subroutine

End Module LionboltHIP
