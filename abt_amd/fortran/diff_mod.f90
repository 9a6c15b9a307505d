! diff_mod.f90 -- drop-in replacement of PBT's far-field diffraction module.
!
! Same module name and the same public procedures, with identical dummy
! arguments, as the reference module (src/diff_mod.f90 of hballington12/abt):
!   diff_main            (reference diff_mod.f90:674-703)
!   diffraction          (reference diff_mod.f90:446-507)
!   make_far_field_bins  (reference diff_mod.f90:827-855)
! The work is done on the GPU by libabt_diff.so through its C-ABI
! (include/abt_diff.h); this module only flattens the outbeamtype and
! geometry_type records (types_mod.f90:8-17, :156-190) into the C structures.
!
! Precision: the environment variable ABT_PRECISION selects "fp64" (default,
! the reference's formulas in FP64), "fp64_fast" (the exact form factor in a
! restructured FP64 body) or "mixed" (FP32 body, FP64 phases).
! Errors from the library stop the program with its message, as the reference
! stops on fatal conditions.
module diff_mod

    use types_mod
    use omp_lib
    use, intrinsic :: iso_c_binding

    implicit none
    private
    public :: diff_main, diffraction, make_far_field_bins

    type, bind(C) :: abt_beams_c
        integer(c_int64_t) :: n
        type(c_ptr) :: ampl, vk7, prop, fov, vert_off, verts
    end type

    type, bind(C) :: abt_grid_c
        integer(c_int64_t) :: n_theta, n_phi
        type(c_ptr) :: theta, phi
    end type

    interface
        integer(c_int) function abt_diff_main_c(grid, lambda, is_fast_diff, beams, ext, precision, &
                                                ab, ae) bind(C, name="abt_diff_main")
            import :: c_int, c_double, c_ptr, abt_grid_c, abt_beams_c
            type(abt_grid_c), intent(in) :: grid
            real(c_double), value :: lambda
            integer(c_int), value :: is_fast_diff
            type(abt_beams_c), intent(in) :: beams, ext
            integer(c_int), value :: precision
            type(c_ptr), value :: ab, ae
        end function

        integer(c_int) function abt_diffraction_c(n_phi, n_theta, x, y, z, lambda, phi, theta, &
                                                  is_fast_diff, ampl, vk7, prop, fov, nv, verts, &
                                                  precision, amplc) bind(C, name="abt_diffraction")
            import :: c_int, c_int64_t, c_double, c_ptr
            integer(c_int64_t), value :: n_phi, n_theta
            type(c_ptr), value :: x, y, z
            real(c_double), value :: lambda
            type(c_ptr), value :: phi, theta
            integer(c_int), value :: is_fast_diff
            type(c_ptr), value :: ampl, vk7, prop
            real(c_double), value :: fov
            integer(c_int64_t), value :: nv
            type(c_ptr), value :: verts
            integer(c_int), value :: precision
            type(c_ptr), value :: amplc
        end function

        integer(c_int) function abt_make_far_field_bins_c(n_theta, theta, n_phi, phi, x, y, z) &
                                                          bind(C, name="abt_make_far_field_bins")
            import :: c_int, c_int64_t, c_ptr
            integer(c_int64_t), value :: n_theta, n_phi
            type(c_ptr), value :: theta, phi, x, y, z
        end function

        type(c_ptr) function abt_last_error_c() bind(C, name="abt_last_error")
            import :: c_ptr
        end function
    end interface

    contains

    subroutine check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer i
        if (rc == 0) return
        call c_f_pointer(abt_last_error_c(), msg, [512])
        text = ' '
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        print*, 'diff_mod (abt_diff): ', where, ' failed: ', trim(text)
        error stop 'diffraction library error'
    end subroutine

    integer(c_int) function precision_flag()
        character(len=16) :: val
        integer :: length, status
        call get_environment_variable("ABT_PRECISION", val, length, status)
        precision_flag = 0
        if (status /= 0) return
        if (val(1:5) == 'mixed') precision_flag = 1
        if (val(1:9) == 'fp64_fast') precision_flag = 3
    end function

    ! Flatten a beam tree: amplitude matrices, e-perp, propagation vectors,
    ! fields of view and the vertices of the facet each beam came from
    ! (geometry%v(geometry%f(fi)%vi(:),:), reference diff_mod.f90:529-531).
    subroutine pack_beams(tree, n, geometry, ampl, vk7, prop, fov, voff, verts)
        type(outbeamtype), dimension(:), intent(in) :: tree
        integer(8), intent(in) :: n
        type(geometry_type), intent(in) :: geometry
        real(8), dimension(:), allocatable, intent(out) :: ampl, vk7, prop, fov, verts
        integer(c_int64_t), dimension(:), allocatable, intent(out) :: voff
        integer(8) :: j, k, fi, nv, total, m
        complex(8) :: a(2, 2)
        allocate(ampl(max(8 * n, 1_8)), vk7(max(3 * n, 1_8)), prop(max(3 * n, 1_8)), &
                 fov(max(n, 1_8)), voff(n + 1))
        total = 0
        voff(1) = 0
        do j = 1, n
            total = total + geometry%f(tree(j)%fi)%nv
            voff(j + 1) = total
        end do
        allocate(verts(max(3 * total, 1_8)))
        m = 0
        do j = 1, n
            a = tree(j)%ampl
            ampl(8 * (j - 1) + 1) = real(a(1, 1)); ampl(8 * (j - 1) + 2) = aimag(a(1, 1))
            ampl(8 * (j - 1) + 3) = real(a(2, 1)); ampl(8 * (j - 1) + 4) = aimag(a(2, 1))
            ampl(8 * (j - 1) + 5) = real(a(1, 2)); ampl(8 * (j - 1) + 6) = aimag(a(1, 2))
            ampl(8 * (j - 1) + 7) = real(a(2, 2)); ampl(8 * (j - 1) + 8) = aimag(a(2, 2))
            vk7(3 * (j - 1) + 1:3 * j) = tree(j)%vk7(:)
            prop(3 * (j - 1) + 1:3 * j) = tree(j)%prop_out(:)
            fov(j) = tree(j)%fov
            fi = tree(j)%fi
            nv = geometry%f(fi)%nv
            do k = 1, nv
                verts(3 * m + 1:3 * m + 3) = geometry%v(geometry%f(fi)%vi(k), :)
                m = m + 1
            end do
        end do
    end subroutine

    subroutine make_far_field_bins(x, y, z, theta_vals, phi_vals)
        real(8), dimension(:,:), allocatable, intent(out) :: x, y, z
        real(8), dimension(:), allocatable, intent(in) :: theta_vals
        real(8), dimension(:), allocatable, intent(in) :: phi_vals
        real(8), dimension(:), allocatable, target :: th, ph
        real(8), dimension(:,:), allocatable, target :: xt, yt, zt
        integer(c_int64_t) :: nth, nph
        nph = size(phi_vals, 1)
        nth = size(theta_vals, 1)
        allocate(xt(nph, nth), yt(nph, nth), zt(nph, nth))
        th = theta_vals
        ph = phi_vals
        call check(abt_make_far_field_bins_c(nth, c_loc(th), nph, c_loc(ph), c_loc(xt), c_loc(yt), &
                                             c_loc(zt)), 'make_far_field_bins')
        call move_alloc(xt, x)
        call move_alloc(yt, y)
        call move_alloc(zt, z)
    end subroutine

    subroutine diffraction(xfar, yfar, zfar, lambda, amplC, phi_vals, theta_vals, job_params, outbeam, &
                           geometry)
        real(8), dimension(:,:), allocatable, intent(in) :: xfar, yfar, zfar
        real(8), intent(in) :: lambda
        complex(8), dimension(:,:,:,:), allocatable, intent(inout) :: amplC
        real(8), dimension(:), allocatable, intent(in) :: phi_vals
        real(8), dimension(:), allocatable, intent(in) :: theta_vals
        type(job_parameters_type), intent(in) :: job_params
        type(outbeamtype), intent(in) :: outbeam
        type(geometry_type), intent(in) :: geometry

        real(8), dimension(:,:), allocatable, target :: x, y, z
        real(8), dimension(:), allocatable, target :: ph, th, verts
        real(8), target :: ampl(8), vk7(3), prop(3)
        complex(8), dimension(:,:,:,:), allocatable, target :: out
        integer(c_int64_t) :: nph, nth, nv
        integer(8) :: k, fi
        integer(c_int) :: fast
        complex(8) :: a(2, 2)

        nph = size(xfar, 1)
        nth = size(xfar, 2)
        x = xfar; y = yfar; z = zfar
        ph = phi_vals; th = theta_vals
        a = outbeam%ampl
        ampl = [real(a(1, 1)), aimag(a(1, 1)), real(a(2, 1)), aimag(a(2, 1)), &
                real(a(1, 2)), aimag(a(1, 2)), real(a(2, 2)), aimag(a(2, 2))]
        vk7 = outbeam%vk7
        prop = outbeam%prop_out
        fi = outbeam%fi
        nv = geometry%f(fi)%nv
        allocate(verts(3 * nv))
        do k = 1, nv
            verts(3 * (k - 1) + 1:3 * k) = geometry%v(geometry%f(fi)%vi(k), :)
        end do
        fast = 0
        if (job_params%is_fast_diff) fast = 1
        allocate(out(nph, nth, 2, 2))
        call check(abt_diffraction_c(nph, nth, c_loc(x), c_loc(y), c_loc(z), lambda, c_loc(ph), &
                                     c_loc(th), fast, c_loc(ampl), c_loc(vk7), c_loc(prop), &
                                     outbeam%fov, nv, c_loc(verts), precision_flag(), c_loc(out)), &
                   'diffraction')
        if (.not. allocated(amplC)) allocate(amplC(nph, nth, 2, 2))
        amplC = out
    end subroutine

    subroutine diff_main(beam_outbeam_tree, beam_outbeam_tree_counter, ampl_far_beam, &
                         ext_diff_outbeam_tree, ampl_far_ext_diff, job_params, geometry)
        type(outbeamtype), dimension(:), allocatable, intent(inout) :: beam_outbeam_tree
        integer(8), intent(inout) :: beam_outbeam_tree_counter
        complex(8), dimension(:,:,:,:), allocatable, intent(out) :: ampl_far_beam
        type(outbeamtype), dimension(:), allocatable, intent(in) :: ext_diff_outbeam_tree
        complex(8), dimension(:,:,:,:), allocatable, intent(out) :: ampl_far_ext_diff
        type(job_parameters_type), intent(in) :: job_params
        type(geometry_type), intent(in) :: geometry

        real(8), dimension(:), allocatable, target :: th, ph
        real(8), dimension(:), allocatable, target :: ba, bk, bp, bf, bv, ea, ek, ep, ef, ev
        integer(c_int64_t), dimension(:), allocatable, target :: bo, eo
        complex(8), dimension(:,:,:,:), allocatable, target :: ab, ae
        type(abt_grid_c) :: grid
        type(abt_beams_c) :: beams, ext, none
        integer(8) :: nph, nth, ne
        integer(c_int) :: fast, prec
        integer(c_int64_t), target :: zoff(1)
        real(8) :: start, finish, start1

        if (job_params%timing) start = omp_get_wtime()
        th = job_params%theta_vals
        ph = job_params%phi_vals
        nth = size(th, 1)
        nph = size(ph, 1)
        ne = 0
        if (allocated(ext_diff_outbeam_tree)) ne = size(ext_diff_outbeam_tree, 1)
        call pack_beams(beam_outbeam_tree, beam_outbeam_tree_counter, geometry, ba, bk, bp, bf, bo, bv)
        if (ne > 0) then
            call pack_beams(ext_diff_outbeam_tree, ne, geometry, ea, ek, ep, ef, eo, ev)
        end if
        allocate(ab(nph, nth, 2, 2), ae(nph, nth, 2, 2))
        grid = abt_grid_c(nth, nph, c_loc(th), c_loc(ph))
        beams = abt_beams_c(beam_outbeam_tree_counter, c_loc(ba), c_loc(bk), c_loc(bp), c_loc(bf), &
                            c_loc(bo), c_loc(bv))
        zoff = 0
        none = abt_beams_c(0_c_int64_t, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(zoff), &
                           c_null_ptr)
        if (ne > 0) then
            ext = abt_beams_c(ne, c_loc(ea), c_loc(ek), c_loc(ep), c_loc(ef), c_loc(eo), c_loc(ev))
        else
            ext = none
        end if
        fast = 0
        if (job_params%is_fast_diff) fast = 1
        prec = precision_flag()

        if (job_params%debug >= 1) write(101,*)'start diff beam loop...'
        if (job_params%timing) start1 = omp_get_wtime()
        ! one tree per call, so that each gets the reference's own timing line
        call check(abt_diff_main_c(grid, job_params%la, fast, beams, none, prec, c_loc(ab), &
                                   c_null_ptr), 'diff_main (beam tree)')
        if (job_params%debug >= 1) then
            write(101,*)'end diff beam loop...'
            if (job_params%timing) then
                finish = omp_get_wtime()
                write(101,'(A,f16.8,A)')"beam diffraction took: ",finish-start1," secs"
                start1 = omp_get_wtime()
            end if
            write(101,*)'start ext diff loop...'
        end if
        ! ext-diff beams take fov = pi/2 inside the library (reference diff_mod.f90:797)
        call check(abt_diff_main_c(grid, job_params%la, fast, none, ext, prec, c_null_ptr, &
                                   c_loc(ae)), 'diff_main (ext diff tree)')
        if (job_params%debug >= 1) then
            write(101,*)'end ext diff loop...'
            if (job_params%timing) then
                finish = omp_get_wtime()
                write(101,'(A,f16.8,A)')"external diffraction took: ",finish-start1," secs"
            end if
        end if
        call move_alloc(ab, ampl_far_beam)
        call move_alloc(ae, ampl_far_ext_diff)
        if (job_params%debug >= 1) then
            if (job_params%timing) then
                finish = omp_get_wtime()
                write(101,'(A,f16.8,A)')"diffraction took: ",finish-start," secs"
            end if
        end if
    end subroutine

end module diff_mod
