NAME          INFEAS_LP
* x >= 5 and x <= 2 cannot both hold.
ROWS
 N  OBJ
 G  C1
 L  C2
COLUMNS
    X         OBJ       1.0   C1        1.0
    X         C2        1.0
RHS
    RHS1      C1        5.0   C2        2.0
ENDATA
