NAME          BOUNDS_ONLY
ROWS
 N  OBJ
 L  C1
COLUMNS
    X1        OBJ       1.0   C1        2.0
RHS
    RHS1      C1        5.0
BOUNDS
 LO BND1      X1        0.0
 LO BND1      X_UNSEEN  3.0
 UP BND1      X_UNSEEN  8.0
ENDATA