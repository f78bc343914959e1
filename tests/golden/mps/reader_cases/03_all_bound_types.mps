NAME          ALL_BOUNDS
ROWS
 N  OBJ
 L  C1
COLUMNS
    X_LO      OBJ       1.0   C1        1.0
    X_UP      OBJ       2.0   C1        1.0
    X_FX      OBJ       3.0   C1        1.0
    X_FR      OBJ       4.0   C1        1.0
RHS
    RHS1      C1        100.0
BOUNDS
 LO BND1      X_LO      2.5
 UP BND1      X_UP      10.0
 FX BND1      X_FX      7.0
 FR BND1      X_FR
ENDATA