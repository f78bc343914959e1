NAME          UNKNOWNSEC
ROWS
 N  OBJ
 L  C1
COLUMNS
    X         OBJ       1.0   C1        1.0
RHS
    RHS1      C1        10.0
NOTASECTION
    RNG1      C1        4.0
ENDATA
