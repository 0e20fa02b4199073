! A Fortran client of the drop-in API through module SimintFortran
! (include/simint/simint.f90): the same job file and result format as
! tests/c_api/client.c.
!
!   client_f <jobs.txt> <out.bin>
program client_f
    use iso_c_binding
    use SimintFortran
    implicit none
    character(len=4096) :: jobf, outf
    integer :: nshell, njob, s, j, p, am, nprim, kind, a, b, c, d, rc, maxam, u, v
    double precision :: x, y, z
    double precision, allocatable :: al(:), co(:), work(:), res(:), cx(:), cy(:), cz(:), cq(:)
    type(c_simint_shell), allocatable, target :: sh(:)
    type(c_simint_multi_shellpair), target :: PP, QQ
    integer(C_INT64_T) :: n
    integer(C_INT32_T) :: rc32

    call get_command_argument(1, jobf)
    call get_command_argument(2, outf)
    call simint_init()
    open(newunit=u, file=trim(jobf), status="old", action="read")
    read(u, *) nshell
    allocate(sh(nshell), cx(nshell), cy(nshell), cz(nshell), cq(nshell))
    maxam = 0
    do s = 1, nshell
        ! "am nprim x y z alpha(1..nprim) coef(1..nprim)" on one line
        read(u, *) am, nprim
        backspace(u)
        allocate(al(nprim), co(nprim))
        read(u, *) am, nprim, x, y, z, (al(p), p = 1, nprim), (co(p), p = 1, nprim)
        call simint_initialize_shell(sh(s))
        call simint_create_shell(nprim, am, x, y, z, al, co, sh(s))
        deallocate(al, co)
        cx(s) = x
        cy(s) = y
        cz(s) = z
        cq(s) = 1.0d0
        maxam = max(maxam, am)
    end do
    call simint_normalize_shells(nshell, sh)
    allocate(work(max(simint_eri_worksize(1, maxam), 1)))
    read(u, *) njob
    open(newunit=v, file=trim(outf), access="stream", form="unformatted", status="replace")
    do j = 1, njob
        read(u, *) kind, a, b, c, d
        ! shell indices in the file count from 0
        a = a + 1
        b = b + 1
        c = c + 1
        d = d + 1
        if (kind <= 1) then
            n = int(ncart(sh(a)%am) * ncart(sh(b)%am) * ncart(sh(c)%am) * ncart(sh(d)%am), C_INT64_T)
            if (kind == 1) n = n * 12
            allocate(res(n))
            call simint_initialize_multi_shellpair(PP)
            call simint_initialize_multi_shellpair(QQ)
            call simint_create_multi_shellpair(1, sh(a:a), 1, sh(b:b), PP, 0)
            call simint_create_multi_shellpair(1, sh(c:c), 1, sh(d:d), QQ, 0)
            if (kind == 1) then
                rc = simint_compute_eri_deriv(1, PP, QQ, 0.0d0, work, res)
            else
                rc = simint_compute_eri(PP, QQ, 0.0d0, work, res)
            end if
            call simint_free_multi_shellpair(PP)
            call simint_free_multi_shellpair(QQ)
        else
            n = int(ncart(sh(a)%am) * ncart(sh(b)%am), C_INT64_T)
            allocate(res(n))
            if (kind == 2) then
                rc = simint_compute_overlap(sh(a), sh(b), res)
            else if (kind == 3) then
                rc = simint_compute_ke(sh(a), sh(b), res)
            else
                rc = simint_compute_potential(nshell, cq, cx, cy, cz, sh(a), sh(b), res)
            end if
        end if
        rc32 = int(rc, C_INT32_T)
        write(v) rc32, n, res
        deallocate(res)
    end do
    close(v)
    close(u)
    call simint_free_shells(nshell, sh)
    call simint_finalize()

contains

    integer function ncart(l)
        integer, intent(in) :: l
        ncart = (l + 1) * (l + 2) / 2
    end function

end program client_f
