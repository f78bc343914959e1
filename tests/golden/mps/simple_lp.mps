NAME          SIMPLE_LP
ROWS
 N  OBJ
 G  C1
COLUMNS
    X1        OBJ       2.0   C1        1.0
    X2        OBJ       3.0   C1        1.0
RHS
    RHS1      C1        4.0
BOUNDS
 UP BND       X1        10.0
 UP BND       X2        10.0
ENDATA
