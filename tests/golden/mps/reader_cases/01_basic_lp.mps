NAME          BASIC_LP
ROWS
 N  OBJ
 L  C_LESS
 G  C_GREATER
 E  C_EQUAL
COLUMNS
    X1        OBJ       3.0   C_LESS    1.0
    X1        C_GREATER 2.0
    X2        OBJ       4.0   C_LESS    2.0
    X2        C_EQUAL   1.0
RHS
    RHS1      C_LESS    10.0  C_GREATER 5.0
    RHS1      C_EQUAL   7.0
ENDATA