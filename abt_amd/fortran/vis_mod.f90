! vis_mod.f90 -- drop-in replacement of the facet-visibility subroutines of PBT's
! near-field tracer (module beam_loop_mod, src/beam_loop_mod.f90 of hballington12/abt):
!   find_vis_int  (reference beam_loop_mod.f90:2037-2247)
!   find_vis_ext  (reference beam_loop_mod.f90:2249-2459)
!   find_vis_inc  (reference beam_loop_mod.f90:2461-2609)
! with identical dummy arguments.  beam_loop_mod drops its own three bodies and
! uses this module; its callers (:487, :676, :876) are unchanged.  The work is done
! on the GPU by libabt_diff.so (include/abt_vis.h, abt_find_vis); this module only
! flattens the geometry_type and beam_type fields those routines read
! (types_mod.f90:100-190) into the C structure.  Results are bit-identical to the
! reference's.  Errors from the library stop the program with its message.
module vis_mod

    use types_mod
    use, intrinsic :: iso_c_binding

    implicit none
    private
    public :: find_vis_int, find_vis_ext, find_vis_inc

    type, bind(C) :: abt_vis_geometry_c
        integer(c_int64_t) :: nv
        type(c_ptr) :: v
        integer(c_int64_t) :: nf
        type(c_ptr) :: f_off, f_vi, f_mid, f_nz, f_ap
        integer(c_int64_t) :: na
        type(c_ptr) :: ap_nz
    end type

    interface
        integer(c_int) function abt_find_vis_c(mode, geo, nf_in, field_in, beam_mid_z, in_beam, is_vis, &
                                               dist_beam, id_beam, is_shad) bind(C, name="abt_find_vis")
            import :: c_int, c_int64_t, c_double, c_ptr, abt_vis_geometry_c
            integer(c_int), value :: mode
            type(abt_vis_geometry_c), intent(in) :: geo
            integer(c_int64_t), value :: nf_in
            type(c_ptr), value :: field_in
            real(c_double), value :: beam_mid_z
            type(c_ptr), value :: in_beam, is_vis, dist_beam, id_beam, is_shad
        end function

        type(c_ptr) function abt_last_error_c() bind(C, name="abt_last_error")
            import :: c_ptr
        end function
    end interface

    contains

    subroutine find_vis_int(in_beam, dist_beam, id_beam, is_shad, beam, rot_geometry)
        logical, dimension(:), allocatable, intent(out) :: in_beam
        real(8), dimension(:), allocatable, intent(out) :: dist_beam
        integer(8), dimension(:), allocatable, intent(out) :: id_beam
        logical, dimension(:), allocatable, intent(out) :: is_shad
        type(beam_type), intent(in) :: beam
        type(geometry_type), intent(in) :: rot_geometry
        logical, dimension(:), allocatable :: is_vis
        call run_vis(0, in_beam, is_vis, dist_beam, id_beam, is_shad, beam, rot_geometry, 0d0)
    end subroutine

    subroutine find_vis_ext(in_beam, dist_beam, id_beam, is_shad, beam, rot_geometry)
        logical, dimension(:), allocatable, intent(out) :: in_beam
        real(8), dimension(:), allocatable, intent(out) :: dist_beam
        integer(8), dimension(:), allocatable, intent(out) :: id_beam
        logical, dimension(:), allocatable, intent(out) :: is_shad
        type(beam_type), intent(in) :: beam
        type(geometry_type), intent(in) :: rot_geometry
        logical, dimension(:), allocatable :: is_vis
        call run_vis(1, in_beam, is_vis, dist_beam, id_beam, is_shad, beam, rot_geometry, 0d0)
    end subroutine

    subroutine find_vis_inc(in_beam, is_vis, dist_beam, id_beam, is_shad, beam, geometry, beam_geometry)
        logical, dimension(:), allocatable, intent(out) :: in_beam
        real(8), dimension(:), allocatable, intent(out) :: dist_beam
        integer(8), dimension(:), allocatable, intent(out) :: id_beam
        logical, dimension(:), allocatable, intent(out) :: is_shad
        logical, dimension(:), allocatable, intent(out) :: is_vis
        type(beam_type), intent(in) :: beam
        type(geometry_type), intent(in) :: geometry
        type(geometry_type), intent(in) :: beam_geometry
        call run_vis(2, in_beam, is_vis, dist_beam, id_beam, is_shad, beam, geometry, beam_geometry%f(1)%mid(3))
    end subroutine

    subroutine run_vis(mode, in_beam, is_vis, dist_beam, id_beam, is_shad, beam, g, beam_mid_z)
        integer, intent(in) :: mode
        logical, dimension(:), allocatable, intent(out) :: in_beam, is_vis, is_shad
        real(8), dimension(:), allocatable, intent(out) :: dist_beam
        integer(8), dimension(:), allocatable, intent(out) :: id_beam
        type(beam_type), intent(in) :: beam
        type(geometry_type), intent(in) :: g
        real(8), intent(in) :: beam_mid_z
        integer(c_int64_t), dimension(:), allocatable, target :: f_off, f_vi, f_ap, fin, idb
        real(c_double), dimension(:,:), allocatable, target :: v, mid
        real(c_double), dimension(:), allocatable, target :: nz, apnz, dist
        integer(c_int8_t), dimension(:), allocatable, target :: ib, iv, sh
        type(abt_vis_geometry_c) :: geo
        integer(8) :: i, nf, nfin
        integer(c_int) :: rc

        nf = g%nf
        allocate(f_off(nf + 1), f_ap(nf), mid(3, nf), nz(nf), apnz(g%na))
        f_off(1) = 0
        do i = 1, nf
            f_off(i + 1) = f_off(i) + g%f(i)%nv
        end do
        allocate(f_vi(f_off(nf + 1)))
        do i = 1, nf
            f_vi(f_off(i) + 1:f_off(i + 1)) = g%f(i)%vi(1:g%f(i)%nv)
            f_ap(i) = g%f(i)%ap
            mid(:, i) = g%f(i)%mid
            nz(i) = g%n(g%f(i)%ni, 3)
        end do
        do i = 1, g%na
            apnz(i) = g%ap(i)%n(3)
        end do
        allocate(v(g%nv, 3))
        v = g%v(1:g%nv, 1:3)
        nfin = 0
        if (mode /= 2) nfin = beam%nf_in
        allocate(fin(max(nfin, 1_8)))
        fin = 0
        do i = 1, nfin
            fin(i) = beam%field_in(i)%fi
        end do
        allocate(ib(nf), iv(nf), sh(nf), dist(nf), idb(nf))
        geo%nv = g%nv
        geo%v = c_loc(v)
        geo%nf = nf
        geo%f_off = c_loc(f_off)
        geo%f_vi = c_loc(f_vi)
        geo%f_mid = c_loc(mid)
        geo%f_nz = c_loc(nz)
        geo%f_ap = c_loc(f_ap)
        geo%na = g%na
        geo%ap_nz = c_loc(apnz)
        rc = abt_find_vis_c(int(mode, c_int), geo, nfin, c_loc(fin), beam_mid_z, c_loc(ib), c_loc(iv), &
                            c_loc(dist), c_loc(idb), c_loc(sh))
        call check(rc)
        allocate(in_beam(nf), is_vis(nf), is_shad(nf), dist_beam(nf), id_beam(nf))
        in_beam = ib /= 0
        is_vis = iv /= 0
        is_shad = sh /= 0
        dist_beam = dist
        id_beam = idb
    end subroutine

    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        character(kind=c_char), dimension(:), pointer :: msg
        character(len=512) :: text
        integer :: k
        if (rc == 0) return
        call c_f_pointer(abt_last_error_c(), msg, [512])
        text = ''
        do k = 1, 512
            if (msg(k) == c_null_char) exit
            text(k:k) = msg(k)
        end do
        print *, 'find_vis: ', trim(text)
        error stop 'abt visibility library error'
    end subroutine

end module vis_mod
