NAME          UNBND_LP
* min -x with x >= 1 and no upper bound.
ROWS
 N  OBJ
 G  C1
COLUMNS
    X         OBJ       -1.0  C1        1.0
RHS
    RHS1      C1        1.0
ENDATA
