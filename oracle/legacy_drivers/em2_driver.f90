! Test infrastructure (oracle/_ref only): runs electromag_cpml_2d from
! legacy/fortran/emFDTD2D.f95 (unmodified) in the current directory.
! usage: legacy_em2 nx nz dx dz npml sx sz nstep
subroutine write_image(image_data, nx, nz, it, channel)
    implicit none
    integer :: nx, nz, it
    real(kind=8) :: image_data(nx, nz)
    character(len=2) :: channel
    character(len=64) :: fname
    write (fname, "(a2, a1, i6.6, a4)") channel, '.', it, '.dat'
    open(unit=10, form='unformatted', file=trim(fname))
    write(10) image_data
    close(unit=10)
end subroutine write_image

program legacy_em2_driver
    implicit none
    interface
        subroutine electromag_cpml_2d(nx, nz, dx, dz, npoints_pml, src, nstep)
            integer :: nx, nz, npoints_pml, nstep
            real(kind=8) :: dx, dz
            integer, dimension(:) :: src
        end subroutine electromag_cpml_2d
    end interface
    integer :: nx, nz, npml, nstep, src(2)
    real(kind=8) :: dx, dz
    character(len=64) :: s
    call get_command_argument(1, s); read(s, *) nx
    call get_command_argument(2, s); read(s, *) nz
    call get_command_argument(3, s); read(s, *) dx
    call get_command_argument(4, s); read(s, *) dz
    call get_command_argument(5, s); read(s, *) npml
    call get_command_argument(6, s); read(s, *) src(1)
    call get_command_argument(7, s); read(s, *) src(2)
    call get_command_argument(8, s); read(s, *) nstep
    call electromag_cpml_2d(nx, nz, dx, dz, npml, src, nstep)
end program legacy_em2_driver
