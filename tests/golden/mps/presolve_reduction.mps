NAME          PRESOLVE_RED
ROWS
 N  OBJ
 G  C1
COLUMNS
    X1        OBJ       2.0   C1        1.0
    X2        OBJ       3.0   C1        1.0
    X3        OBJ       4.0   C1        1.0
RHS
    RHS1      C1        9.0
BOUNDS
 UP BND       X1        10.0
 UP BND       X2        10.0
 FX BND       X3        5.0
ENDATA
