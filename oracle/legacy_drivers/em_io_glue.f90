! Test infrastructure (oracle/_ref only).  legacy/fortran/emFDTD2D.f95 and
! emFDTD25D.f95 hold the same arithmetic as electromag2 / electromag25 in
! legacy/fortran/electromagfdtd.f95 (that file does not compile: its complex
! variant calls maxval on a COMPLEX array).  They call their file I/O as
! external procedures; these are that I/O, in the record layouts of
! legacy/fortran/readwrite_routines.f95 (material_rw :451-470, loadsource
! :9-23, loadcpml :26-37).  Snapshots are FP64 records named {ch}.{it:06}.dat.

subroutine material_rw(filename, image_data, readfile)
    implicit none
    character(len=*) :: filename
    real(kind=8) :: image_data(*)
    logical :: readfile
    integer(kind=8) :: nbytes
    integer :: n
    inquire(file=trim(filename), size=nbytes)
    n = int((nbytes - 8) / 8)
    open(unit=13, form='unformatted', file=trim(filename))
    if (readfile) then
        read(13) image_data(1:n)
    end if
    close(unit=13)
end subroutine material_rw

subroutine loadsource(filename, n, srcfn)
    implicit none
    character(len=*) :: filename
    integer :: n
    real(kind=8) :: srcfn(n)
    open(unit=13, form='unformatted', file=trim(filename))
    read(13) srcfn
    close(unit=13)
end subroutine loadsource

subroutine loadcpml(filename, image_data)
    implicit none
    character(len=*) :: filename
    real(kind=8) :: image_data(*)
    integer(kind=8) :: nbytes
    integer :: n
    inquire(file=trim(filename), size=nbytes)
    n = int(nbytes / 8)
    open(unit=13, form='unformatted', file=trim(filename), access='stream')
    read(13) image_data(1:n)
    close(unit=13)
end subroutine loadcpml
