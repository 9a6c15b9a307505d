! Test driver of the drop-in diff_mod: reads a beam set written by
! tests/fortran/__init__.py, fills the reference's own data types and calls
! diff_main, diffraction and make_far_field_bins exactly as PBT's drivers do
! (pbt_seq_mod.f90:212-218, beam_loop_mod.f90:44, :1865), then writes the
! results.  Usage: abt_driver <input.bin> <output.bin> <log.txt>; ABT_DRIVER_REPEAT=n calls
! diff_main n times (the reference drivers' once per orientation), each writing its timing lines.
program abt_driver
    use types_mod
    use diff_mod
    implicit none
    character(len=512) :: fin, fout, flog
    integer(8) :: hdr(7), nth, nph, nvert, nface, nb, ne, j, k, off
    real(8) :: lam
    type(job_parameters_type) :: job_params
    type(geometry_type) :: geometry
    type(outbeamtype), dimension(:), allocatable :: beams, ext
    complex(8), dimension(:,:,:,:), allocatable :: ampl_far_beam, ampl_far_ext_diff, amplc
    real(8), dimension(:,:), allocatable :: x, y, z
    integer(8), dimension(:), allocatable :: fnv, fvi
    real(8), dimension(:), allocatable :: buf
    character(len=16) :: rep_s
    integer :: rep, ios, r

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    call get_command_argument(3, flog)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) hdr
    nth = hdr(1); nph = hdr(2); nvert = hdr(3); nface = hdr(4); nb = hdr(5); ne = hdr(6)
    read(10) lam
    job_params%la = lam
    job_params%is_fast_diff = hdr(7) /= 0
    job_params%is_multithreaded = .false.
    job_params%timing = .true.
    job_params%debug = 1
    allocate(job_params%theta_vals(nth), job_params%phi_vals(nph))
    read(10) job_params%theta_vals
    read(10) job_params%phi_vals
    allocate(geometry%v(nvert, 3))
    read(10) geometry%v
    allocate(fnv(nface))
    read(10) fnv
    allocate(fvi(sum(fnv)))
    read(10) fvi
    geometry%nv = nvert
    geometry%nf = nface
    allocate(geometry%f(nface))
    off = 0
    do j = 1, nface
        geometry%f(j)%nv = fnv(j)
        allocate(geometry%f(j)%vi(fnv(j)))
        geometry%f(j)%vi = fvi(off + 1:off + fnv(j))
        off = off + fnv(j)
    end do
    call read_tree(beams, nb)
    call read_tree(ext, ne)
    close(10)

    rep = 1
    call get_environment_variable('ABT_DRIVER_REPEAT', rep_s, status=ios)
    if (ios == 0) read(rep_s, *, iostat=ios) rep
    if (ios /= 0 .or. rep < 1) rep = 1
    open(101, file=trim(flog))
    do r = 1, rep
        call diff_main(beams, nb, ampl_far_beam, ext, ampl_far_ext_diff, job_params, geometry)
    end do
    call make_far_field_bins(x, y, z, job_params%theta_vals, job_params%phi_vals)
    allocate(amplc(nph, nth, 2, 2))
    if (nb > 0) then
        call diffraction(x, y, z, job_params%la, amplc, job_params%phi_vals, job_params%theta_vals, &
                         job_params, beams(1), geometry)
    else
        amplc = 0
    end if
    close(101)

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) ampl_far_beam
    write(11) ampl_far_ext_diff
    write(11) x
    write(11) y
    write(11) z
    write(11) amplc
    close(11)

contains

    subroutine read_tree(tree, n)
        type(outbeamtype), dimension(:), allocatable, intent(out) :: tree
        integer(8), intent(in) :: n
        real(8), allocatable :: a(:), v(:), p(:), f(:)
        integer(8), allocatable :: fi(:)
        integer(8) :: i
        allocate(tree(n), a(8 * n), v(3 * n), p(3 * n), f(n), fi(n))
        if (n == 0) return
        read(10) a
        read(10) v
        read(10) p
        read(10) f
        read(10) fi
        do i = 1, n
            tree(i)%ampl(1, 1) = cmplx(a(8 * i - 7), a(8 * i - 6), 8)
            tree(i)%ampl(2, 1) = cmplx(a(8 * i - 5), a(8 * i - 4), 8)
            tree(i)%ampl(1, 2) = cmplx(a(8 * i - 3), a(8 * i - 2), 8)
            tree(i)%ampl(2, 2) = cmplx(a(8 * i - 1), a(8 * i), 8)
            tree(i)%vk7 = v(3 * i - 2:3 * i)
            tree(i)%prop_out = p(3 * i - 2:3 * i)
            tree(i)%prop_in = p(3 * i - 2:3 * i)
            tree(i)%fov = f(i)
            tree(i)%fi = fi(i)
        end do
    end subroutine

end program abt_driver
