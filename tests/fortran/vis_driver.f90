! Test driver of the drop-in vis_mod: reads a geometry and an illuminating facet list
! written by tests/test_gpu_fortran.py, fills the reference's own geometry_type and
! beam_type, calls find_vis_int, find_vis_ext and find_vis_inc as the tracer does
! (beam_loop_mod.f90:487, :676, :876) and writes their outputs.
! Usage: abt_vis_driver <input.bin> <output.bin>
program abt_vis_driver
    use types_mod
    use vis_mod
    implicit none
    character(len=512) :: fin, fout
    integer(8) :: hdr(4), nv, nf, na, nin, j, off
    type(geometry_type) :: geometry, beam_geometry
    type(beam_type) :: beam
    integer(8), dimension(:), allocatable :: fnv, fvi, fap, fi
    real(8), dimension(:,:), allocatable :: mid, apn
    real(8) :: bz
    logical, dimension(:), allocatable :: in_beam, is_vis, is_shad
    real(8), dimension(:), allocatable :: dist_beam
    integer(8), dimension(:), allocatable :: id_beam

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) hdr
    nv = hdr(1); nf = hdr(2); na = hdr(3); nin = hdr(4)
    geometry%nv = nv
    geometry%nf = nf
    geometry%na = na
    allocate(geometry%v(nv, 3), geometry%n(nf, 3), fnv(nf))
    read(10) geometry%v
    read(10) fnv
    allocate(fvi(sum(fnv)), mid(3, nf), fap(nf), apn(na, 3), fi(nin))
    read(10) fvi
    read(10) mid
    read(10) geometry%n
    read(10) fap
    read(10) apn
    read(10) fi
    read(10) bz
    close(10)
    allocate(geometry%f(nf), geometry%ap(na))
    off = 0
    do j = 1, nf
        geometry%f(j)%nv = fnv(j)
        allocate(geometry%f(j)%vi(fnv(j)))
        geometry%f(j)%vi = fvi(off + 1:off + fnv(j))
        off = off + fnv(j)
        geometry%f(j)%mid = mid(:, j)
        geometry%f(j)%ni = j
        geometry%f(j)%ap = fap(j)
    end do
    do j = 1, na
        geometry%ap(j)%n = apn(j, :)
    end do
    beam%nf_in = nin
    allocate(beam%field_in(nin))
    do j = 1, nin
        beam%field_in(j)%fi = fi(j)
    end do
    allocate(beam_geometry%f(1))
    beam_geometry%f(1)%mid = [0d0, 0d0, bz]

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    call find_vis_int(in_beam, dist_beam, id_beam, is_shad, beam, geometry)
    call put(.false.)
    call find_vis_ext(in_beam, dist_beam, id_beam, is_shad, beam, geometry)
    call put(.false.)
    call find_vis_inc(in_beam, is_vis, dist_beam, id_beam, is_shad, beam, geometry, beam_geometry)
    call put(.true.)
    close(11)

contains

    subroutine put(with_vis)
        logical, intent(in) :: with_vis
        integer(1), dimension(:), allocatable :: b
        allocate(b(nf))
        b = merge(1_1, 0_1, in_beam)
        write(11) b
        b = 0_1
        if (with_vis) b = merge(1_1, 0_1, is_vis)
        write(11) b
        write(11) dist_beam
        write(11) id_beam
        b = merge(1_1, 0_1, is_shad)
        write(11) b
    end subroutine

end program abt_vis_driver
