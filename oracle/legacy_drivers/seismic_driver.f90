! Test infrastructure (oracle/_ref only): runs the legacy seismic2 / seismic25
! subroutines, compiled unmodified from legacy/fortran/seismicfdtd.f95 and
! legacy/fortran/readwrite_routines.f95, in the current directory.  Inputs are
! the legacy file contract (Fortran records for coefficients and sources,
! stream files for the 1-D CPML profiles); snapshots are written as FP64
! records by the legacy write_image2/3 (OUTPUT_SINGLE = .false.).
!
! usage: legacy_seismic 2  nx nz dx dz npml sx sz nstep
!        legacy_seismic 25 nx ny nz dx dy dz npml sx sy sz nstep
program legacy_seismic_driver
    use seismicfdtd
    implicit none
    integer :: kind, nx, ny, nz, npml, nstep
    real(kind=8) :: dx, dy, dz
    integer :: src2(2), src3(3)

    kind = iarg(1)
    ! seismic2 never initialises its gamma_z array (seismicfdtd.f95:158-159
    ! read gammas_z.dat into gamma_x); zero the stack its frame will occupy
    ! so that array reads 0 (no attenuation on vz) instead of garbage.
    call zero_stack(512 * 1024)
    if (kind == 2) then
        nx = iarg(2); nz = iarg(3)
        dx = rarg(4); dz = rarg(5)
        npml = iarg(6)
        src2(1) = iarg(7); src2(2) = iarg(8)
        nstep = iarg(9)
        call seismic2(nx, nz, dx, dz, npml, src2, nstep, .false.)
    else
        nx = iarg(2); ny = iarg(3); nz = iarg(4)
        dx = rarg(5); dy = rarg(6); dz = rarg(7)
        npml = iarg(8)
        src3(1) = iarg(9); src3(2) = iarg(10); src3(3) = iarg(11)
        nstep = iarg(12)
        call seismic25(nx, ny, nz, dx, dy, dz, npml, src3, nstep, .false.)
    end if

contains

    subroutine zero_stack(n)
        integer, intent(in) :: n
        real(kind=8) :: scratch(n)
        integer :: q
        do q = 1, n
            scratch(q) = 0.d0
        end do
        call keep(scratch, n)
    end subroutine zero_stack

    subroutine keep(a, n)
        integer, intent(in) :: n
        real(kind=8), intent(in) :: a(n)
        if (sum(a) /= 0.d0) print *, a(1)
    end subroutine keep

    integer function iarg(n)
        integer, intent(in) :: n
        character(len=64) :: s
        call get_command_argument(n, s)
        read(s, *) iarg
    end function iarg

    real(kind=8) function rarg(n)
        integer, intent(in) :: n
        character(len=64) :: s
        call get_command_argument(n, s)
        read(s, *) rarg
    end function rarg

end program legacy_seismic_driver
