! Fortran binding of the simint drop-in API (module SimintFortran, the same
! module, type and procedure names as skel/simint/simint.f90, so Fortran
! callers of the reference switch over unchanged).
!
! The two derived types mirror struct simint_shell and struct
! simint_multi_shellpair of include/simint/shell/shell.h; this library always
! has the derivative fields alpha2/beta2 (its SIMINT_OSTEI_MAXDER is 1).
! Every procedure is a thin wrapper over the C entry point of the same name:
! default-kind arguments are converted to C kinds and arrays and structures
! are passed by address.  Compile with any Fortran 2003 compiler and link
! against libsimint.so.
module SimintFortran
    use iso_c_binding
    implicit none
    private

    public :: c_simint_shell, c_simint_multi_shellpair
    public :: simint_init, simint_finalize
    public :: simint_initialize_shell, simint_initialize_shells, simint_allocate_shell
    public :: simint_free_shell, simint_free_shells, simint_normalize_shells, simint_copy_shell
    public :: simint_create_shell, simint_create_zero_shell
    public :: simint_initialize_multi_shellpair, simint_initialize_multi_shellpairs
    public :: simint_free_multi_shellpair, simint_free_multi_shellpairs
    public :: simint_create_multi_shellpair, simint_create_multi_shellpair2
    public :: simint_compute_eri, simint_compute_eri_deriv
    public :: simint_eri_worksize, simint_eri_workmem
    public :: simint_compute_overlap, simint_compute_ke, simint_compute_potential

    ! struct simint_shell
    type, bind(C) :: c_simint_shell
        integer(C_INT) :: am, nprim
        real(C_DOUBLE) :: x, y, z
        type(C_PTR) :: alpha, coef
        integer(C_SIZE_T) :: memsize
        type(C_PTR) :: ptr
    end type c_simint_shell

    ! struct simint_multi_shellpair
    type, bind(C) :: c_simint_multi_shellpair
        integer(C_INT) :: am1, am2, nprim, nshell12, nshell12_clip
        type(C_PTR) :: nprim12
        type(C_PTR) :: AB_x, AB_y, AB_z
        type(C_PTR) :: x, y, z
        type(C_PTR) :: PA_x, PA_y, PA_z
        type(C_PTR) :: PB_x, PB_y, PB_z
        type(C_PTR) :: alpha, alpha2, beta2
        type(C_PTR) :: prefac, screen
        real(C_DOUBLE) :: screen_max
        integer(C_SIZE_T) :: memsize
        type(C_PTR) :: ptr
    end type c_simint_multi_shellpair

    ! the C entry points (include/simint/*.h)
    interface
        subroutine simint_init() bind(C, name="simint_init")
        end subroutine
        subroutine simint_finalize() bind(C, name="simint_finalize")
        end subroutine
        subroutine cs_initialize_shell(G) bind(C, name="simint_initialize_shell")
            import :: C_PTR
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_initialize_shells(n, G) bind(C, name="simint_initialize_shells")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_allocate_shell(nprim, G) bind(C, name="simint_allocate_shell")
            import :: C_PTR, C_INT
            integer(C_INT), value :: nprim
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_free_shell(G) bind(C, name="simint_free_shell")
            import :: C_PTR
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_free_shells(n, G) bind(C, name="simint_free_shells")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_normalize_shells(n, G) bind(C, name="simint_normalize_shells")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_copy_shell(src, dest) bind(C, name="simint_copy_shell")
            import :: C_PTR
            type(C_PTR), value :: src, dest
        end subroutine
        subroutine cs_create_shell(nprim, am, x, y, z, alpha, coef, G) bind(C, name="simint_create_shell")
            import :: C_PTR, C_INT, C_DOUBLE
            integer(C_INT), value :: nprim, am
            real(C_DOUBLE), value :: x, y, z
            type(C_PTR), value :: alpha, coef, G
        end subroutine
        subroutine cs_create_zero_shell(G) bind(C, name="simint_create_zero_shell")
            import :: C_PTR
            type(C_PTR), value :: G
        end subroutine
        subroutine cs_initialize_multi_shellpair(P) bind(C, name="simint_initialize_multi_shellpair")
            import :: C_PTR
            type(C_PTR), value :: P
        end subroutine
        subroutine cs_initialize_multi_shellpairs(n, P) bind(C, name="simint_initialize_multi_shellpairs")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: P
        end subroutine
        subroutine cs_free_multi_shellpair(P) bind(C, name="simint_free_multi_shellpair")
            import :: C_PTR
            type(C_PTR), value :: P
        end subroutine
        subroutine cs_free_multi_shellpairs(n, P) bind(C, name="simint_free_multi_shellpairs")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: P
        end subroutine
        subroutine cs_create_multi_shellpair(na, A, nb, B, P, screen_method) &
                bind(C, name="simint_create_multi_shellpair")
            import :: C_PTR, C_INT
            integer(C_INT), value :: na, nb, screen_method
            type(C_PTR), value :: A, B, P
        end subroutine
        subroutine cs_create_multi_shellpair2(npair, AB, P, screen_method) &
                bind(C, name="simint_create_multi_shellpair2")
            import :: C_PTR, C_INT
            integer(C_INT), value :: npair, screen_method
            type(C_PTR), value :: AB, P
        end subroutine
        integer(C_INT) function cs_compute_eri(P, Q, screen_tol, work, integrals) &
                bind(C, name="simint_compute_eri")
            import :: C_PTR, C_INT, C_DOUBLE
            type(C_PTR), value :: P, Q, work, integrals
            real(C_DOUBLE), value :: screen_tol
        end function
        integer(C_INT) function cs_compute_eri_deriv(deriv, P, Q, screen_tol, work, integrals) &
                bind(C, name="simint_compute_eri_deriv")
            import :: C_PTR, C_INT, C_DOUBLE
            integer(C_INT), value :: deriv
            type(C_PTR), value :: P, Q, work, integrals
            real(C_DOUBLE), value :: screen_tol
        end function
        integer(C_SIZE_T) function cs_eri_worksize(derorder, maxam) bind(C, name="simint_eri_worksize")
            import :: C_INT, C_SIZE_T
            integer(C_INT), value :: derorder, maxam
        end function
        integer(C_SIZE_T) function cs_eri_workmem(derorder, maxam) bind(C, name="simint_eri_workmem")
            import :: C_INT, C_SIZE_T
            integer(C_INT), value :: derorder, maxam
        end function
        integer(C_INT) function cs_compute_overlap(a, b, integrals) bind(C, name="simint_compute_overlap")
            import :: C_PTR, C_INT
            type(C_PTR), value :: a, b, integrals
        end function
        integer(C_INT) function cs_compute_ke(a, b, integrals) bind(C, name="simint_compute_ke")
            import :: C_PTR, C_INT
            type(C_PTR), value :: a, b, integrals
        end function
        integer(C_INT) function cs_compute_potential(n, charge, cx, cy, cz, a, b, integrals) &
                bind(C, name="simint_compute_potential")
            import :: C_PTR, C_INT
            integer(C_INT), value :: n
            type(C_PTR), value :: charge, cx, cy, cz, a, b, integrals
        end function
    end interface

contains

    subroutine simint_initialize_shell(G)
        type(c_simint_shell), intent(inout), target :: G
        call cs_initialize_shell(c_loc(G))
    end subroutine

    subroutine simint_initialize_shells(n, G)
        integer, intent(in) :: n
        type(c_simint_shell), intent(inout), target :: G(n)
        call cs_initialize_shells(int(n, C_INT), c_loc(G))
    end subroutine

    subroutine simint_allocate_shell(nprim, G)
        integer, intent(in) :: nprim
        type(c_simint_shell), intent(inout), target :: G
        call cs_allocate_shell(int(nprim, C_INT), c_loc(G))
    end subroutine

    subroutine simint_free_shell(G)
        type(c_simint_shell), intent(inout), target :: G
        call cs_free_shell(c_loc(G))
    end subroutine

    subroutine simint_free_shells(n, G)
        integer, intent(in) :: n
        type(c_simint_shell), intent(inout), target :: G(n)
        call cs_free_shells(int(n, C_INT), c_loc(G))
    end subroutine

    subroutine simint_normalize_shells(n, G)
        integer, intent(in) :: n
        type(c_simint_shell), intent(inout), target :: G(n)
        call cs_normalize_shells(int(n, C_INT), c_loc(G))
    end subroutine

    subroutine simint_copy_shell(src, dest)
        type(c_simint_shell), intent(in), target :: src
        type(c_simint_shell), intent(inout), target :: dest
        call cs_copy_shell(c_loc(src), c_loc(dest))
    end subroutine

    subroutine simint_create_shell(nprim, am, x, y, z, alpha, coef, G)
        integer, intent(in) :: nprim, am
        double precision, intent(in) :: x, y, z
        double precision, intent(in), target :: alpha(nprim), coef(nprim)
        type(c_simint_shell), intent(inout), target :: G
        call cs_create_shell(int(nprim, C_INT), int(am, C_INT), real(x, C_DOUBLE), real(y, C_DOUBLE), &
                             real(z, C_DOUBLE), c_loc(alpha), c_loc(coef), c_loc(G))
    end subroutine

    subroutine simint_create_zero_shell(G)
        type(c_simint_shell), intent(inout), target :: G
        call cs_create_zero_shell(c_loc(G))
    end subroutine

    subroutine simint_initialize_multi_shellpair(P)
        type(c_simint_multi_shellpair), intent(inout), target :: P
        call cs_initialize_multi_shellpair(c_loc(P))
    end subroutine

    subroutine simint_initialize_multi_shellpairs(n, P)
        integer, intent(in) :: n
        type(c_simint_multi_shellpair), intent(inout), target :: P(n)
        call cs_initialize_multi_shellpairs(int(n, C_INT), c_loc(P))
    end subroutine

    subroutine simint_free_multi_shellpair(P)
        type(c_simint_multi_shellpair), intent(inout), target :: P
        call cs_free_multi_shellpair(c_loc(P))
    end subroutine

    subroutine simint_free_multi_shellpairs(n, P)
        integer, intent(in) :: n
        type(c_simint_multi_shellpair), intent(inout), target :: P(n)
        call cs_free_multi_shellpairs(int(n, C_INT), c_loc(P))
    end subroutine

    subroutine simint_create_multi_shellpair(na, A, nb, B, P, screen_method)
        integer, intent(in) :: na, nb, screen_method
        type(c_simint_shell), intent(in), target :: A(na), B(nb)
        type(c_simint_multi_shellpair), intent(inout), target :: P
        call cs_create_multi_shellpair(int(na, C_INT), c_loc(A), int(nb, C_INT), c_loc(B), c_loc(P), &
                                       int(screen_method, C_INT))
    end subroutine

    subroutine simint_create_multi_shellpair2(npair, AB, P, screen_method)
        integer, intent(in) :: npair, screen_method
        type(c_simint_shell), intent(in), target :: AB(2 * npair)
        type(c_simint_multi_shellpair), intent(inout), target :: P
        call cs_create_multi_shellpair2(int(npair, C_INT), c_loc(AB), c_loc(P), int(screen_method, C_INT))
    end subroutine

    ! number of shell quartets computed, -1 if screened out, or a negative
    ! SIMINT_ERROR_* code
    integer function simint_compute_eri(P, Q, screen_tol, work, integrals) result(res)
        type(c_simint_multi_shellpair), intent(in), target :: P, Q
        real(C_DOUBLE), intent(in) :: screen_tol
        double precision, intent(inout), target :: work(*), integrals(*)
        res = int(cs_compute_eri(c_loc(P), c_loc(Q), screen_tol, c_loc(work), c_loc(integrals)))
    end function

    integer function simint_compute_eri_deriv(deriv, P, Q, screen_tol, work, integrals) result(res)
        integer, intent(in) :: deriv
        type(c_simint_multi_shellpair), intent(in), target :: P, Q
        real(C_DOUBLE), intent(in) :: screen_tol
        double precision, intent(inout), target :: work(*), integrals(*)
        res = int(cs_compute_eri_deriv(int(deriv, C_INT), c_loc(P), c_loc(Q), screen_tol, c_loc(work), &
                                       c_loc(integrals)))
    end function

    integer function simint_eri_worksize(derorder, maxam) result(res)
        integer, intent(in) :: derorder, maxam
        res = int(cs_eri_worksize(int(derorder, C_INT), int(maxam, C_INT)))
    end function

    integer function simint_eri_workmem(derorder, maxam) result(res)
        integer, intent(in) :: derorder, maxam
        res = int(cs_eri_workmem(int(derorder, C_INT), int(maxam, C_INT)))
    end function

    integer function simint_compute_overlap(P, Q, integrals) result(res)
        type(c_simint_shell), intent(in), target :: P, Q
        double precision, intent(inout), target :: integrals(*)
        res = int(cs_compute_overlap(c_loc(P), c_loc(Q), c_loc(integrals)))
    end function

    integer function simint_compute_ke(P, Q, integrals) result(res)
        type(c_simint_shell), intent(in), target :: P, Q
        double precision, intent(inout), target :: integrals(*)
        res = int(cs_compute_ke(c_loc(P), c_loc(Q), c_loc(integrals)))
    end function

    integer function simint_compute_potential(n, nuc_charge, X, Y, Z, P, Q, integrals) result(res)
        integer, intent(in) :: n
        double precision, intent(in), target :: nuc_charge(n), X(n), Y(n), Z(n)
        type(c_simint_shell), intent(in), target :: P, Q
        double precision, intent(inout), target :: integrals(*)
        res = int(cs_compute_potential(int(n, C_INT), c_loc(nuc_charge), c_loc(X), c_loc(Y), c_loc(Z), &
                                       c_loc(P), c_loc(Q), c_loc(integrals)))
    end function

end module SimintFortran
