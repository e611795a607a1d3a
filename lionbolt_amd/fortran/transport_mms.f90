!  Driver in the shape of validation/operators/TransportMMS.f90:47-89, running on
!  the GPU through the LionboltHIP shim.
!
!  Usage: transport_mms <input.bin> <output.bin>
!  input (stream, native endian): NE, NK, NI, L (int64); rg(3,NK), k(3,NI), w(NI)
!  (real64); node(4,NE), neighbors(4,NE), mat(NE) (int64, 1-based, boundary 0);
!  sigma (real64), c (real64: scattering ratio of the Boltzmann solve).
!  output: T^-1 s for the MMS source (NENK*NI real64), then the scalar flux of
!  BoltzmannOp%MatInv of a unit source (NENK real64), then the first-collision
!  source of a planar beam along (theta, phi) = (0.41, 2.2), a direction no ray
!  follows along a mesh edge (ExternalBeam%Source with that operator's physics,
!  NENK*NI real64), and its uncollided angular fluence for attenuation 0.5
!  (ExternalBeam%AddAngularFluence, NENK*NI real64).  Prints the RMS MMS error
!  and the error of the delta-down kernel in the shape of
!  validation/operators/deltadownMMS.f90 (XSdd = 1, 10 on two materials, s = 1).
Program transport_mms
    use iso_c_binding
    use LionboltHIP
    Implicit None
    Type (HIPOperator)               :: OpT, OpL, OpK
    Integer (c_int64_t)              :: NE, NK, NI, L, e, i, kk, sdof, n
    Real (c_double),     Allocatable :: rg (:,:), kv (:,:), w (:), s (:), strue (:), phi (:), sigs (:,:)
    Real (c_double),     Allocatable :: sb (:), ub (:)
    Type (lb_beam_desc)              :: beam
    Integer (c_int64_t), Allocatable :: node (:,:), nbr (:,:), mat (:), mat2 (:)
    Real (c_double)                  :: sigma, c, r (3), err
    Real (c_double), Parameter       :: A = 10.0_c_double
    Character (LEN=512)              :: fin, fout
    Integer                          :: u

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
    read (u) NE, NK, NI, L
    ALLOCATE(rg(3, NK), kv(3, NI), w(NI), node(4, NE), nbr(4, NE), mat(NE))
    read (u) rg, kv, w, node, nbr, mat, sigma, c
    close (u)

    ! TransportMMS: psi = exp(-A r^2), source (-2A k.r + sigma) exp(-A r^2)
    call OpT%BuildSpaceAngle(rg, node, nbr, mat, kv, w, INT(L), 'GMRES')
    ALLOCATE(sigs(0:L, 1))
    sigs = 0.0_c_double
    call OpT%BuildPhysics([sigma], sigs)
    n = 4 * NE * NI
    ALLOCATE(s(n), strue(n))
    do i = 1, NI
        do e = 1, NE
            do kk = 1, 4
                sdof = 4 * (e - 1) + kk
                r = rg(:, node(kk, e))
                s(sdof + (i - 1) * 4 * NE) = (-2.0_c_double * A * DOT_PRODUCT(kv(:, i), r) + sigma) &
                                             * EXP(-A * DOT_PRODUCT(r, r))
                strue(sdof + (i - 1) * 4 * NE) = EXP(-A * DOT_PRODUCT(r, r))
            end do
        end do
    end do
    call OpT%TransportMatInv(s)
    err = SQRT(SUM((s - strue)**2) / n)
    write (*, '(A,ES12.4)') 'TransportMMS RMS error: ', err

    ! BoltzmannOp%MatInv of a unit isotropic source with scattering ratio c
    call OpL%BuildSpaceAngle(rg, node, nbr, mat, kv, w, INT(L), 'GMRES')
    sigs = 0.0_c_double
    sigs(0, 1) = c
    call OpL%BuildPhysics([1.0_c_double], sigs)
    strue = 1.0_c_double
    call OpL%BoltzmannMatInv(strue)
    phi = OpL%Fluence(strue)
    write (*, '(A,I0,A,ES12.4)') 'BoltzmannOp%MatInv iterates: ', OpL%stats%iterates, &
        '  true residual: ', OpL%stats%true_residual

    ! A planar beam through the shim: its first-collision source with OpL's
    ! physics, and its uncollided angular fluence for attenuation 0.5
    beam%axis = [SIN(0.41_c_double) * COS(2.2_c_double), SIN(0.41_c_double) * SIN(2.2_c_double), COS(0.41_c_double)]
    beam%origin = -3.0_c_double * beam%axis
    beam%e1 = [1.0_c_double, 0.0_c_double, 0.0_c_double] - beam%axis(1) * beam%axis
    beam%e1 = beam%e1 / NORM2(beam%e1)
    beam%e2 = [beam%axis(2) * beam%e1(3) - beam%axis(3) * beam%e1(2), &
               beam%axis(3) * beam%e1(1) - beam%axis(1) * beam%e1(3), &
               beam%axis(1) * beam%e1(2) - beam%axis(2) * beam%e1(1)]
    beam%cutout = LB_CUTOUT_NONE
    beam%params = 0.0_c_double
    beam%n = 1.0_c_double
    ALLOCATE(sb(n), ub(n))
    sb = 0.0_c_double
    call OpL%BeamSource(beam, sb)
    ub = 0.0_c_double
    call OpL%BeamAddAngularFluence(beam, [0.5_c_double], ub)
    write (*, '(A,ES12.4)') 'ExternalBeam%Source sum: ', SUM(sb)

    ! deltadownMMS: two materials split at z = 0, K = XSdd(mat) I, s = 1
    ALLOCATE(mat2(NE))
    do e = 1, NE
        mat2(e) = MERGE(2_c_int64_t, 1_c_int64_t, SUM(rg(3, node(:, e))) > 0.0_c_double)
    end do
    call OpK%BuildSpaceAngle(rg, node, nbr, mat2, kv, w, INT(L), 'GMRES')
    call OpK%BuildDeltaDown([1.0_c_double, 10.0_c_double])
    strue = 1.0_c_double
    call OpK%ScatterMatVec(strue)
    err = 0.0_c_double
    do i = 1, NI
        do e = 1, NE
            do kk = 1, 4
                sdof = 4 * (e - 1) + kk
                err = MAX(err, ABS(strue(sdof + (i - 1) * 4 * NE) - MERGE(10.0_c_double, 1.0_c_double, mat2(e) == 2)))
            end do
        end do
    end do
    write (*, '(A,ES12.4)') 'deltadownMMS max error: ', err
    call OpK%Destroy()

    open (newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
    write (u) s, phi, sb, ub
    close (u)
    call OpT%Destroy()
    call OpL%Destroy()
    print *, 'transport_mms - Successfully terminated.'
End Program
