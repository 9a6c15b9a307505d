! Stand-in for PBT's types_mod (src/types_mod.f90 of hballington12/abt) for the test
! drivers: only the derived-type components that the drop-in modules
! (abt_amd/fortran/diff_mod.f90, vis_mod.f90) and the drivers touch, with the
! component names and kinds PBT declares, so the same module sources compile
! against PBT's own types_mod in a PBT build and against this one here.  Test
! infrastructure: it lets the drivers build without the PBT sources.
module types_mod
    implicit none

    ! outgoing beam of a beam tree (types_mod.f90:8-17)
    type outbeamtype
        complex(8) ampl(1:2, 1:2)
        real(8) vk7(1:3)
        real(8) prop_out(1:3)
        real(8) prop_in(1:3)
        integer(8) fi
        real(8) fov
    end type outbeamtype

    ! run options the far-field path reads
    type job_parameters_type
        real(8) la
        logical is_multithreaded
        real(8), dimension(:), allocatable :: theta_vals
        real(8), dimension(:), allocatable :: phi_vals
        integer(8) debug
        logical timing
        logical is_fast_diff
    end type job_parameters_type

    ! illuminating facets of a near-field beam
    type field_in_type
        integer(8) fi
    end type field_in_type

    type beam_type
        type(field_in_type), dimension(:), allocatable :: field_in
        integer(8) nf_in
    end type beam_type

    type facet_type
        integer(8), dimension(:), allocatable :: vi
        integer(8) ni
        real(8) mid(1:3)
        integer(8) nv
        integer(8) ap
    end type facet_type

    type aperture_type
        real(8) n(1:3)
    end type aperture_type

    type geometry_type
        real(8), dimension(:, :), allocatable :: v
        real(8), dimension(:, :), allocatable :: n
        type(facet_type), dimension(:), allocatable :: f
        integer(8) nv
        integer(8) nf
        integer(8) na
        type(aperture_type), dimension(:), allocatable :: ap
    end type geometry_type
end module types_mod
