* This is a header comment line
NAME          NEG_COMMENTS

ROWS
 N  OBJ_ROW
 L  CON1

* Comment in COLUMNS section
COLUMNS

    X1        OBJ_ROW   -5.25   CON1      -3.0
    X2        OBJ_ROW   10.0    CON1      1.5

* Blank lines and comments before RHS
RHS
    RHS1      CON1      -15.0

ENDATA