! ISO_C_BINDING interfaces to libuegccd (include/uegccd.h).  The derived
! types are bind(C) mirrors of the C structs, field for field.
Module Wrappers
  Use, Intrinsic :: ISO_C_Binding
  Implicit None

  Integer(c_int), Parameter :: UEGCCD_OK = 0, UEGCCD_ERR_NOT_CONVERGED = -5
  Integer(c_int32_t), Parameter :: UEGCCD_MODE_REFERENCE = 0, UEGCCD_MODE_DIIS = 1
  Integer(c_int32_t), Parameter :: UEGCCD_PREC_FP64 = 0, UEGCCD_PREC_MIXED = 1

  Type, Bind(C) :: uegccd_system
    Integer(c_int32_t) :: nocc, nao
    Type(c_ptr) :: nvec, eigen
    Real(c_double) :: L, madelung, rs, gapii
    Integer(c_int32_t) :: coren, safe_eri
    Integer(c_int32_t) :: irange(15)
    Integer(c_int32_t) :: do_ring, do_xring, do_ladder, do_mosaic
    Real(c_double) :: screening
  End Type uegccd_system

  Type, Bind(C) :: uegccd_options
    Integer(c_int32_t) :: mode, ndiis, diis_start, max_iter
    Real(c_double) :: tol_dt, tol_e, tol_r, damping
    Integer(c_int32_t) :: precision
    Real(c_double) :: mixed_switch
    Integer(c_int32_t) :: device, iters_per_sync, graph, persistent
  End Type uegccd_options

  Type, Bind(C) :: uegccd_result
    Real(c_double) :: ecorr, ehf_plus_ec, dt, maxres
    Integer(c_int32_t) :: niter, failed, iters_fp32, status
    Real(c_double) :: solve_ms
  End Type uegccd_result

  Type, Bind(C) :: uegccd_heg_params
    Integer(c_int32_t) :: nel, maxkpoint
    Real(c_double) :: rs, twist(3), fcut, corefcut, madelungfactor, gap, exchangefactor
    Integer(c_int32_t) :: cutoff_twisted = 0   ! 0: the reference's |n|^2 cutoff
  End Type uegccd_heg_params

  Type, Bind(C) :: uegccd_server_options
    Integer(c_int32_t) :: device
    Integer(c_int64_t) :: hbm_budget_bytes
    Integer(c_int32_t) :: max_active, iters_per_step, planner_threads
  End Type uegccd_server_options

  Type, Bind(C) :: uegccd_request
    Type(uegccd_heg_params) :: heg
    Real(c_double) :: gapii
    Integer(c_int32_t) :: coren, safe_eri
    Integer(c_int32_t) :: irange(15)
    Integer(c_int32_t) :: do_ring, do_xring, do_ladder, do_mosaic
    Type(uegccd_options) :: opt
  End Type uegccd_request

  Type, Bind(C) :: uegccd_request_result
    Integer(c_int32_t) :: state, status, nao, nocc, niter, iters_fp32
    Real(c_double) :: ehf, e2, ecorr, maxres, t_submit_ms, t_start_ms, t_done_ms
    Integer(c_int64_t) :: footprint_bytes
    Real(c_double) :: plan_ms
  End Type uegccd_request_result

  Interface
    Subroutine uegccd_default_options(opt) Bind(C, name="uegccd_default_options")
      Import :: uegccd_options
      Type(uegccd_options), Intent(Out) :: opt
    End Subroutine

    Integer(c_int) Function uegccd_heg_build(p, nao, nvec, eigen, scalars) &
        Bind(C, name="uegccd_heg_build")
      Import :: uegccd_heg_params, c_int, c_int32_t, c_ptr
      Type(uegccd_heg_params), Intent(In) :: p
      Integer(c_int32_t), Intent(Out) :: nao
      Type(c_ptr), Value :: nvec, eigen, scalars
    End Function

    Integer(c_int) Function uegccd_heg_build_device(p, device, nao, nvec, eigen, scalars) &
        Bind(C, name="uegccd_heg_build_device")
      Import :: uegccd_heg_params, c_int, c_int32_t, c_ptr
      Type(uegccd_heg_params), Intent(In) :: p
      Integer(c_int32_t), Value :: device
      Integer(c_int32_t), Intent(Out) :: nao
      Type(c_ptr), Value :: nvec, eigen, scalars
    End Function

    Integer(c_int) Function uegccd_ctx_create(sys, opt, ctx) Bind(C, name="uegccd_ctx_create")
      Import :: uegccd_system, uegccd_options, c_int, c_ptr
      Type(uegccd_system), Intent(In) :: sys
      Type(uegccd_options), Intent(In) :: opt
      Type(c_ptr), Intent(Out) :: ctx
    End Function

    Subroutine uegccd_ctx_destroy(ctx) Bind(C, name="uegccd_ctx_destroy")
      Import :: c_ptr
      Type(c_ptr), Value :: ctx
    End Subroutine

    Integer(c_int) Function uegccd_mp2(ctx, taa, tab, tba, e2) Bind(C, name="uegccd_mp2")
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(Out) :: taa(*), tab(*), tba(*)
      Real(c_double), Intent(Out) :: e2
    End Function

    Integer(c_int) Function uegccd_solve(ctx, taa, tab, tba, res) Bind(C, name="uegccd_solve")
      Import :: c_int, c_ptr, c_double, uegccd_result
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(InOut) :: taa(*), tab(*), tba(*)
      Type(uegccd_result), Intent(Out) :: res
    End Function

    Integer(c_int) Function uegccd_drpa(ctx, t2, res, esosex) Bind(C, name="uegccd_drpa")
      Import :: c_int, c_ptr, c_double, uegccd_result
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(InOut) :: t2(*)
      Type(uegccd_result), Intent(Out) :: res
      Real(c_double), Intent(Out) :: esosex
    End Function

    Integer(c_int) Function uegccd_trace(ctx, e, dt, n) Bind(C, name="uegccd_trace")
      Import :: c_int, c_ptr, c_double, c_int32_t
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(Out) :: e(*), dt(*)
      Integer(c_int32_t), Value :: n
    End Function

    Type(c_ptr) Function uegccd_strerror_c(st) Bind(C, name="uegccd_strerror")
      Import :: c_ptr, c_int
      Integer(c_int), Value :: st
    End Function
    Integer(c_int) Function uegccd_structure_factor(ctx, tab, tba, ncell, qvec, vals, ecorr) &
        Bind(C, name="uegccd_structure_factor")
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: tab(*), tba(*)
      Integer(c_int32_t), Intent(InOut) :: ncell
      Type(c_ptr), Value :: qvec, vals
      Real(c_double), Intent(Out) :: ecorr
    End Function

    Integer(c_int) Function uegccd_mp2_structure_factor(ctx, ncell, qvec, vals, e2) &
        Bind(C, name="uegccd_mp2_structure_factor")
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Intent(InOut) :: ncell
      Type(c_ptr), Value :: qvec, vals
      Real(c_double), Intent(Out) :: e2
    End Function

    Integer(c_int) Function uegccd_exchange_structure_factor(p, ncell, qvec, count, exv) &
        Bind(C, name="uegccd_exchange_structure_factor")
      Import :: c_int, c_int32_t, c_ptr, uegccd_heg_params
      Type(uegccd_heg_params), Intent(In) :: p
      Integer(c_int32_t), Intent(InOut) :: ncell
      Type(c_ptr), Value :: qvec, count, exv
    End Function

    Integer(c_int) Function uegccd_connectivity(ctx, counts, nbins) Bind(C, name="uegccd_connectivity")
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Intent(Out) :: counts(*)
      Integer(c_int32_t), Value :: nbins
    End Function

    Integer(c_int) Function uegccd_server_create(opt, srv) Bind(C, name="uegccd_server_create")
      Import :: c_int, c_ptr, uegccd_server_options
      Type(uegccd_server_options), Intent(In) :: opt
      Type(c_ptr), Intent(Out) :: srv
    End Function

    Subroutine uegccd_server_destroy(srv) Bind(C, name="uegccd_server_destroy")
      Import :: c_ptr
      Type(c_ptr), Value :: srv
    End Subroutine

    Integer(c_int) Function uegccd_server_submit(srv, req, id) Bind(C, name="uegccd_server_submit")
      Import :: c_int, c_ptr, c_int64_t, uegccd_request
      Type(c_ptr), Value :: srv
      Type(uegccd_request), Intent(In) :: req
      Integer(c_int64_t), Intent(Out) :: id
    End Function

    Integer(c_int) Function uegccd_server_drain(srv) Bind(C, name="uegccd_server_drain")
      Import :: c_int, c_ptr
      Type(c_ptr), Value :: srv
    End Function

    Integer(c_int) Function uegccd_server_poll(srv, id, res) Bind(C, name="uegccd_server_poll")
      Import :: c_int, c_ptr, c_int64_t, uegccd_request_result
      Type(c_ptr), Value :: srv
      Integer(c_int64_t), Value :: id
      Type(uegccd_request_result), Intent(Out) :: res
    End Function

  End Interface

Contains

  ! The library's message for a status, as a Fortran string.
  Function uegccd_message(st) Result(msg)
    Integer(c_int), Intent(In) :: st
    Character(Len=80) :: msg
    Character(Kind=c_char), Pointer :: cs(:)
    Integer :: k
    msg = ' '
    Call c_f_pointer(uegccd_strerror_c(st), cs, [80])
    Do k = 1, 80
      If (cs(k) == c_null_char) Exit
      msg(k:k) = cs(k)
    End Do
  End Function uegccd_message

  Subroutine Check(st, what)
    Integer(c_int), Intent(In) :: st
    Character(Len=*), Intent(In) :: what
    If (st /= UEGCCD_OK) Then
      Write(6,'(A,": ",A)') what, Trim(uegccd_message(st))
      Stop 'uegccd library error'
    End If
  End Subroutine Check

End Module Wrappers
