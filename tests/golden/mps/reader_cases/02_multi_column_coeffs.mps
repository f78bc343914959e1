NAME          MULTI_COEFF
ROWS
 N  COST
 L  ROW1
 G  ROW2
COLUMNS
    VAR1      COST      1.5   ROW1      2.5
    VAR2      COST      -0.5  ROW2      4.0
RHS
    RHS1      ROW1      12.0  ROW2      3.0
ENDATA