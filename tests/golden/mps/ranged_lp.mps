NAME          RANGED_LP
* RANGES on an L row: 6 <= x + y <= 10. min x + y -> 6.
ROWS
 N  OBJ
 L  C1
COLUMNS
    X         OBJ       1.0   C1        1.0
    Y         OBJ       1.0   C1        1.0
RHS
    RHS1      C1        10.0
RANGES
    RNG1      C1        4.0
ENDATA
