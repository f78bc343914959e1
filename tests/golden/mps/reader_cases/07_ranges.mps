NAME          RANGES_ALL
* RANGES after RHS: the case that used to be parsed AS RHS data,
* silently overwriting every right-hand side.
ROWS
 N  OBJ
 L  C1
 G  C2
 E  C3
 E  C4
COLUMNS
    X         OBJ       1.0   C1        1.0
    X         C2        1.0   C3        1.0
    X         C4        1.0
RHS
    RHS1      C1        10.0  C2        5.0
    RHS1      C3        7.0   C4        7.0
RANGES
    RNG1      C1        4.0   C2        4.0
    RNG1      C3        3.0   C4        -3.0
ENDATA
